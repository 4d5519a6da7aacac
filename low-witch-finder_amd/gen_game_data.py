"""Writes csrc/gen/whf_game_data.c, the biome search trees, the terrain spline and the game's sine table in the layout
csrc/whf_game_data.h declares, from the committed data file data/whf_game_data.json.xz.

The data are the game's (dumped from each Minecraft version by the reference's tools/accuracy/dump_trees.py, its
TerrainTable export and its JVM dump of Mth.SIN, shipped in the reference as engine/tools/lysh/src/biome_trees.h,
terrain_table.h and sin_table.h under lysh's MIT licence; data/NOTICE.md gives the provenance and the licence). They
have to be identical entry for entry, so they are converted, not retyped: with --ref the script parses the
reference's headers, checks the structure it reads (subtree ends, child counts, spline pool bounds, sine special
values), and writes the data file (--pack) or fails when the committed data file differs from them.

    python gen_game_data.py [--out csrc/gen/whf_game_data.c]          from the committed data file
    python gen_game_data.py --ref /root/reference [--pack]              check (or rewrite) the data file first
"""
import argparse
import json
import lzma
import os
import re
import struct
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
DATA = os.path.join(HERE, "data", "whf_game_data.json.xz")


def _ints(text):
    return [int(v) for v in re.findall(r"-?\d+", text)]


def parse_trees(src):
    trees = []
    for t in range(64):
        m = re.search(r"gtree%d_range\[\]\[12\]\s*=\s*\{(.*?)\n\};" % t, src, re.S)
        if not m:
            break
        rows = re.findall(r"\{([^{}]*)\}", m.group(1))
        ranges = [_ints(r) for r in rows]
        if any(len(r) != 12 for r in ranges):
            raise SystemExit("tree %d: a range row does not hold 12 values" % t)
        arr = {}
        for name in ("kids", "biome", "end"):
            a = re.search(r"gtree%d_%s\[\]\s*=\s*\{([^}]*)\};" % (t, name), src)
            if not a:
                raise SystemExit("tree %d: no %s array" % (t, name))
            arr[name] = _ints(a.group(1))
        n = len(ranges)
        if not all(len(arr[k]) == n for k in arr):
            raise SystemExit("tree %d: array lengths differ" % t)
        depth = check_tree(t, ranges, arr["kids"], arr["end"])
        trees.append((ranges, arr["kids"], arr["biome"], arr["end"], depth))
    if not trees:
        raise SystemExit("no trees found")
    vt = re.search(r"lysh_versions\[\]\s*=\s*\{(.*?)\n\};", src, re.S)
    if not vt:
        raise SystemExit("no version table")
    versions = [(m[0], int(m[1]), int(m[2]), int(m[3]))
                for m in re.findall(r'\{"([^"]+)",\s*(\d+),\s*(\d+),\s*(\d+)\}', vt.group(1))]
    if not versions or any(v[1] >= len(trees) for v in versions):
        raise SystemExit("bad version table")
    return trees, versions


def check_tree(t, ranges, kids, end):
    """Pre-order layout: the children of node i start at i + 1 and follow each other through end[]; the
    subtree of i ends at end[i]. Returns the tree's depth in nodes (root to deepest leaf)."""
    n = len(kids)
    if end[0] != n:
        raise SystemExit("tree %d: root subtree ends at %d, not %d" % (t, end[0], n))
    stack = [(0, 1)]
    seen = 0
    depth = 0
    while stack:
        i, dep = stack.pop()
        seen += 1
        depth = max(depth, dep)
        if kids[i] == 0:
            if end[i] != i + 1:
                raise SystemExit("tree %d: leaf %d end %d" % (t, i, end[i]))
            continue
        c = i + 1
        for _ in range(kids[i]):
            stack.append((c, dep + 1))
            c = end[c]
        if c != end[i]:
            raise SystemExit("tree %d: children of %d end at %d, not %d" % (t, i, c, end[i]))
    if seen != n:
        raise SystemExit("tree %d: %d of %d nodes reachable" % (t, seen, n))
    if max(end) > 0xFFFF or max(kids) > 0xFF:
        raise SystemExit("tree %d: does not fit the node layout" % t)
    if any(r[2 * k] > r[2 * k + 1] for r in ranges for k in range(6)):
        raise SystemExit("tree %d: a range with min > max" % t)
    return depth


def _f32(tok):
    v = float(tok.rstrip("fF"))
    return struct.unpack("<f", struct.pack("<f", v))[0]


def parse_spline(src):
    fl = re.search(r"LYSH_TERRAIN_FLOATS\[\]\s*=\s*\{(.*?)\};", src, re.S)
    ch = re.search(r"LYSH_TERRAIN_CHILD\[\]\s*=\s*\{(.*?)\};", src, re.S)
    nd = re.search(r"LYSH_TERRAIN_NODES\[\]\s*=\s*\{(.*?)\n\};", src, re.S)
    if not (fl and ch and nd):
        raise SystemExit("terrain table does not parse")
    floats = [_f32(x) for x in re.findall(r"-?[0-9.]+(?:[eE][-+]?\d+)?f", fl.group(1))]
    child = _ints(ch.group(1))
    nodes = []
    for row in re.findall(r"\{([^{}]*)\}", nd.group(1)):
        parts = [p.strip() for p in row.split(",")]
        if len(parts) != 7:
            raise SystemExit("spline node row: %r" % row)
        nodes.append(tuple(int(p) for p in parts[:6]) + (_f32(parts[6]),))
    roots = []
    for name in ("OFFSET", "FACTOR", "PEAK"):
        m = re.search(r"#define\s+LYSH_TERRAIN_SURFACE_%s\s+(\d+)" % name, src)
        if not m:
            raise SystemExit("no root %s" % name)
        roots.append(int(m.group(1)))
    nn = re.search(r"#define\s+LYSH_TERRAIN_NODE_N\s+(\d+)", src)
    if not nn or int(nn.group(1)) != len(nodes):
        raise SystemExit("spline node count mismatch")

    def depth(i):
        kind, coord, n, loc, der, c0, _ = nodes[i]
        if kind == 1:
            return 1
        if kind != 0 or not (0 <= coord < 4) or n < 1:
            raise SystemExit("spline node %d: kind %d coord %d n %d" % (i, kind, coord, n))
        if loc + n > len(floats) or der + n > len(floats) or c0 + n > len(child):
            raise SystemExit("spline node %d: pools out of range" % i)
        # the interval search assumes strictly increasing locations, as the game's spline builder requires
        if any(not floats[loc + k] < floats[loc + k + 1] for k in range(n - 1)):
            raise SystemExit("spline node %d: locations not increasing" % i)
        return 1 + max(depth(child[c0 + k]) for k in range(n))

    d = max(depth(r) for r in roots)
    return floats, child, nodes, roots, d


def parse_sin(src):
    m = re.search(r"LYSH_MC_SIN\[65536\]\s*=\s*\{(.*?)\};", src, re.S)
    if not m:
        raise SystemExit("sin table does not parse")
    vals = [float.fromhex(t.rstrip("fF")) for t in re.findall(r"-?0x[0-9a-fA-F.]+p[-+]?\d+f", m.group(1))]
    if len(vals) != 65536:
        raise SystemExit("sin table has %d entries" % len(vals))
    vals = [struct.unpack("<f", struct.pack("<f", v))[0] for v in vals]
    # spot checks of the table's meaning: sin(0) = 0 and sin at a quarter turn = 1
    if vals[0] != 0.0 or vals[16384] != 1.0:
        raise SystemExit("sin table: unexpected values at 0 or a quarter turn")
    return vals


def c_float(v):
    if v == 0.0:
        return "-0.0f" if struct.pack("<f", v) == struct.pack("<f", -0.0) else "0.0f"
    return float.hex(v) + "f"


def emit(trees, versions, spline, sin_table, out):
    floats, child, nodes, roots, depth = spline
    lines = ["/* Generated by gen_game_data.py from data/whf_game_data.json.xz (the game's tables; data/NOTICE.md); "
             "do not edit. */",
             '#include "../whf_game_data.h"', ""]
    base = [0]
    for ranges, kids, biome, end, _ in trees:
        base.append(base[-1] + len(kids))
    lines.append("const uint32_t whf_gd_tree_base[WHF_GD_TREES + 1] = {%s};" % ",".join(map(str, base)))
    if len(trees) != 8:
        raise SystemExit("expected 8 trees, found %d" % len(trees))
    lines.append("const int whf_gd_tree_depth = %d;" % max(t[4] for t in trees))
    lines.append("const whf_gd_node whf_gd_nodes[] = {")
    for ranges, kids, biome, end, _ in trees:
        for i in range(len(kids)):
            lines.append("{{%s},%d,%d,%d,0}," % (",".join(map(str, ranges[i])), end[i], kids[i], biome[i]))
    lines.append("};")
    lines.append("const whf_gd_version whf_gd_versions[] = {")
    for name, tree, is118, pre in versions:
        lines.append('{"%s",%d,%d,%d},' % (name, tree, is118, pre))
    lines.append("};")
    lines.append("const int whf_gd_n_versions = %d;" % len(versions))
    lines.append("const whf_gd_spline_node whf_gd_spline_nodes[] = {")
    for kind, coord, n, loc, der, c0, const in nodes:
        lines.append("{%d,%d,%d,%d,%d,%d,%s}," % (kind, coord, n, loc, der, c0, c_float(const)))
    lines.append("};")
    lines.append("const int whf_gd_n_spline_nodes = %d;" % len(nodes))
    lines.append("const float whf_gd_spline_floats[] = {%s};" % ",".join(c_float(v) for v in floats))
    lines.append("const int whf_gd_n_spline_floats = %d;" % len(floats))
    lines.append("const int32_t whf_gd_spline_child[] = {%s};" % ",".join(map(str, child)))
    lines.append("const int whf_gd_n_spline_child = %d;" % len(child))
    lines.append("const int whf_gd_spline_root[3] = {%s};" % ",".join(map(str, roots)))
    lines.append("const int whf_gd_spline_depth = %d;" % depth)
    lines.append("const float whf_gd_mc_sin[65536] = {")
    for i in range(0, 65536, 16):
        lines.append(",".join(c_float(v) for v in sin_table[i:i + 16]) + ",")
    lines.append("};")
    os.makedirs(os.path.dirname(out), exist_ok=True)
    tmp = out + ".tmp"
    with open(tmp, "w") as f:
        f.write("\n".join(lines) + "\n")
    os.replace(tmp, out)


def _bits(v):
    return struct.unpack("<I", struct.pack("<f", v))[0]


def _float(b):
    return struct.unpack("<f", struct.pack("<I", b))[0]


def to_data(trees, versions, spline, sin_table):
    """The parsed tables as a JSON-able dict; floats as their IEEE-754 single bit patterns, so nothing rounds."""
    floats, child, nodes, roots, depth = spline
    return {
        "format": 1,
        "trees": [{"ranges": [v for r in ranges for v in r], "kids": kids, "biome": biome, "end": end, "depth": d}
                  for ranges, kids, biome, end, d in trees],
        "versions": [list(v) for v in versions],
        "spline": {"floats": [_bits(v) for v in floats], "child": child,
                   "nodes": [list(n[:6]) + [_bits(n[6])] for n in nodes], "roots": roots, "depth": depth},
        "sin": [_bits(v) for v in sin_table],
    }


def from_data(d):
    if d.get("format") != 1:
        raise SystemExit("unknown game data format")
    trees = []
    for t in d["trees"]:
        r = t["ranges"]
        ranges = [r[12 * i:12 * i + 12] for i in range(len(t["kids"]))]
        trees.append((ranges, t["kids"], t["biome"], t["end"], t["depth"]))
    versions = [tuple(v) for v in d["versions"]]
    sp = d["spline"]
    nodes = [tuple(n[:6]) + (_float(n[6]),) for n in sp["nodes"]]
    spline = ([_float(b) for b in sp["floats"]], sp["child"], nodes, sp["roots"], sp["depth"])
    return trees, versions, spline, [_float(b) for b in d["sin"]]


def read_reference(ref):
    src_dir = os.path.join(ref, "engine", "tools", "lysh", "src")
    with open(os.path.join(src_dir, "biome_trees.h")) as f:
        trees, versions = parse_trees(f.read())
    with open(os.path.join(src_dir, "terrain_table.h")) as f:
        spline = parse_spline(f.read())
    with open(os.path.join(src_dir, "sin_table.h")) as f:
        sin_table = parse_sin(f.read())
    return to_data(trees, versions, spline, sin_table)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--ref", help="the reference tree: check the data file against its tables")
    ap.add_argument("--pack", action="store_true", help="with --ref: (re)write the data file from its tables")
    ap.add_argument("--data", default=DATA)
    ap.add_argument("--out", default=os.path.join(HERE, "csrc", "gen", "whf_game_data.c"))
    a = ap.parse_args(argv)
    if a.ref:
        want = read_reference(a.ref)
        if a.pack:
            os.makedirs(os.path.dirname(a.data), exist_ok=True)
            with lzma.open(a.data + ".tmp", "wt", preset=9 | lzma.PRESET_EXTREME) as f:
                json.dump(want, f, separators=(",", ":"))
            os.replace(a.data + ".tmp", a.data)
            print("wrote %s" % a.data)
        else:
            with lzma.open(a.data, "rt") as f:
                if json.load(f) != want:
                    raise SystemExit("%s differs from the reference's tables (run with --pack)" % a.data)
            print("%s equals the reference's tables" % a.data)
    with lzma.open(a.data, "rt") as f:
        trees, versions, spline, sin_table = from_data(json.load(f))
    # the same structural checks as on the reference's headers
    for t, (ranges, kids, biome, end, depth) in enumerate(trees):
        if check_tree(t, ranges, kids, end) != depth:
            raise SystemExit("tree %d: depth mismatch" % t)
    if sin_table[0] != 0.0 or sin_table[16384] != 1.0 or len(sin_table) != 65536:
        raise SystemExit("sin table: unexpected values")
    emit(trees, versions, spline, sin_table, a.out)
    print("game data: %d trees (%d nodes), %d versions, %d spline nodes (depth %d) -> %s"
          % (len(trees), sum(len(t[1]) for t in trees), len(versions), len(spline[2]), spline[4], a.out))


if __name__ == "__main__":
    sys.exit(main())
