"""Parity of the HIP Deep Dive (whf_deepdive / whf_room_floors through the C ABI) with the CPU oracle, the
reference's mstree vectors and its confirmed deep huts (data/finds.tsv, kept in tests/golden/mstree_vectors.json).

Bar: integer work, so everything is exact: start counts, walk outcomes, LCG draw and piece totals, the spot sets
(S, region, chunk, carver seed, room floor, pieces), the expansion's phase-1 funnel and tier histograms and finds.
"""
import json
import os

from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle_ctypes as O
from product import whf

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
M48 = (1 << 48) - 1


@pytest.fixture(scope="module")
def ctx():
    print("opening whf context on device 0", flush=True)
    c = whf.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def mst():
    with open(os.path.join(GOLD, "mstree_vectors.json")) as f:
        return json.load(f)


def sweep_start():
    """First output of SplitMix64(3), low 48 bits: the start of the 2^32-seed Deep Dive workload."""
    return O.splitmix(3, 0) & M48


def random_starts(n, seed):
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        for X in rng.integers(0, 1 << 48, size=4 * n, dtype=np.uint64):
            if O.is_start(int(X)):
                out.append(int(X))
                if len(out) == n:
                    break
    return out


def test_room_floors_match_oracle_and_mstree(ctx, mst):
    carvers = random_starts(3000, 1)
    carvers += [int(x) for x in np.random.default_rng(2).integers(0, 1 << 48, size=1000, dtype=np.uint64)]
    for v in mst["starts"] + [f for f in mst["finds"] if f["start"]]:
        carvers.append(O.carver_seed(v["seed"] & M48, v["cx"], v["cz"]))
    print("walking %d layouts" % len(carvers), flush=True)
    st, fl, pc = ctx.room_floors(carvers)
    for i, X in enumerate(carvers):
        s, f, p, r = O.room_floor(X)
        assert int(st[i]) == s, (i, X)
        assert int(pc[i]) == p, (i, X)
        if s == 0:
            assert int(fl[i]) == f, (i, X)
    gold = mst["starts"] + [f for f in mst["finds"] if f["start"]]
    base = len(carvers) - len(gold)
    for k, v in enumerate(gold):
        assert (int(st[base + k]), int(fl[base + k]), int(pc[base + k])) == (0, v["room_floor"], v["pieces"])


@pytest.mark.parametrize("floor_max,rect,count,first", [
    (-55, (-2, 2, -2, 2), 1 << 16, None),
    (-50, (-3, 1, 0, 2), 1 << 15, (1 << 48) - 5000),
    (-60, (-1, 1, -1, 1), 1 << 16, 123456789),
])
def test_sweep_matches_oracle(ctx, floor_max, rect, count, first):
    s0 = sweep_start() if first is None else first
    print("sweep S0=%d count=%d rect=%s floor<=%d" % (s0, count, rect, floor_max), flush=True)
    spots, finds, st = ctx.deepdive(s0, count, rect, floor_max=floor_max)
    ref = O.deepdive_sweep(s0, count, rect, floor_max=floor_max, cut=-12 - floor_max, threads=8)
    nreg = (rect[1] - rect[0]) * (rect[3] - rect[2])
    assert st["structure_seeds"] == count and st["tests"] == count * nreg
    assert st["starts"] == ref["starts"]
    assert st["walks"] == ref["status"]
    assert st["draws"] == ref["draws"] and st["pieces"] == ref["pieces"]
    assert spots == ref["spots"]
    assert st["spots"] == len(spots)
    assert finds == [] and st["expanded_seeds"] == 0


def test_known_deep_huts_are_found(ctx, mst):
    """Every confirmed deep hut of the reference's data/finds.tsv comes out of the Deep Dive: its structure seed's
    sweep over the hut's region keeps the chunk (room floor <= -55), and the 2^16 expansion with the phase-1 check
    returns the world seed."""
    for f in mst["finds"]:
        rx, rz = f["cx"] >> 5, f["cz"] >> 5
        print("find seed=%d chunk=(%d,%d)" % (f["seed"], f["cx"], f["cz"]), flush=True)
        spots, finds, st = ctx.deepdive(f["seed"] & M48, 1, (rx, rx + 1, rz, rz + 1), floor_max=-55, expand=True,
                                        max_height=-50, biome=False)
        assert [(s[3], s[4], s[6]) for s in spots] == [(f["cx"], f["cz"], f["room_floor"])]
        assert st["expanded_seeds"] == 1 << 16
        assert (f["seed"], rx, rz, f["cx"], f["cz"], f["room_floor"]) in finds
        # they are witch huts in the game, so the swamp check keeps them
        _, finds_b, st_b = ctx.deepdive(f["seed"] & M48, 1, (rx, rx + 1, rz, rz + 1), floor_max=-55, expand=True,
                                        max_height=-50)
        assert (f["seed"], rx, rz, f["cx"], f["cz"], f["room_floor"]) in finds_b
        assert st_b["biome_checked"] == len(finds) and st_b["biome_rejected"] == len(finds) - len(finds_b)


def expansion_reference(spots, max_y=-50, vflags=0):
    """The oracle's phase-1 scan of every spot's 2^16 world seeds at the spot's region, summed."""
    tot = dict(funnel=[0] * 11, tiers=[0] * 7, hits=[])
    floor_of = {(s[0], s[3], s[4]): s[6] for s in spots}
    for s in spots:
        s48, rx, rz = s[0], s[1], s[2]
        seeds = np.array([O.s64((u << 48) | s48) for u in range(1 << 16)], dtype=np.int64)
        r = O.scan(seeds, (rx, rx + 1, rz, rz + 1), max_y=max_y, vflags=vflags, threads=16)
        tot["funnel"] = [a + b for a, b in zip(tot["funnel"], r["funnel"])]
        tot["tiers"] = [a + b for a, b in zip(tot["tiers"], r["tiers"])]
        tot["hits"] += [h + (floor_of[(h[0] & M48, h[3], h[4])],) for h in r["hits"]]
    tot["hits"].sort()
    return tot


def default_tree(vflags):
    """The rule flags' default biome tree, as lysh picks it (eval.c:104-106)."""
    return whf.biome_tree("1.18.2" if vflags & whf.V_118 else ("26.2" if vflags & whf.V_262 else "1.21.5"))


def biome_filter(hits, vflags):
    """The oracle's swamp check of phase-1 survivors (lysh's biome test at the chunk middle's pre-carver surface,
    biome.c:94-156): the survivors it keeps, and how many it rejects."""
    tree = default_tree(vflags)

    def ok(h):
        seed, cx, cz = h[0] & 0xFFFFFFFFFFFFFFFF, h[3], h[4]
        height, _ = O.Terrain(seed, vflags & whf.V_LARGE).height(cx * 16 + 8, cz * 16 + 8)
        return O.Biome(seed, vflags & whf.V_LARGE).check(tree, cx * 16, cz * 16, height - 1)["winner"] == 6

    with ThreadPoolExecutor(16) as ex:
        keep = list(ex.map(ok, hits))
    return [h for h, k in zip(hits, keep) if k], keep.count(False)


@pytest.mark.timeout(600)
def test_expansion_matches_oracle_scan(ctx):
    """Phase-1 chain of the 2^16 expansion equal to the oracle's scan; then the swamp check of the survivors equal to
    the oracle's biome test, the finds being the survivors it keeps."""
    s0 = sweep_start()
    spots, finds, st = ctx.deepdive(s0, 1 << 16, (-2, 2, -2, 2), floor_max=-55, expand=True, max_height=-50)
    assert len(spots) > 16   # more than one expansion batch
    print("expanding %d spots" % len(spots), flush=True)
    ref = expansion_reference(spots)
    assert st["expanded_seeds"] == len(spots) << 16
    assert st["funnel"] == ref["funnel"]
    assert st["tiers"] == ref["tiers"]
    assert st["accepted"] == len(ref["hits"])
    print("biome check of %d survivors" % len(ref["hits"]), flush=True)
    want, rejected = biome_filter(ref["hits"], 0)
    assert st["biome_checked"] == len(ref["hits"]) and st["biome_rejected"] == rejected
    assert finds == want
    # without the check the finds are the phase-1 survivors
    _, finds1, st1 = ctx.deepdive(s0, 1 << 16, (-2, 2, -2, 2), floor_max=-55, expand=True, max_height=-50,
                                  biome=False)
    assert finds1 == ref["hits"] and st1["biome_checked"] == 0


@pytest.mark.timeout(300)
def test_expansion_swamp_check_keeps_and_rejects_as_oracle(ctx):
    """Enough expansion survivors for the swamp check to both keep and reject: a 2^20-seed sweep (518 spots), phase 1
    at maxHeight 60, gives 136 survivors of which the check rejects 6. The survivors come from the same call without
    the check (phase 1's parity with the oracle is the test above); each is put to the oracle's biome test."""
    s0 = sweep_start()
    rect = (-2, 2, -2, 2)
    spots, surv, st0 = ctx.deepdive(s0, 1 << 20, rect, floor_max=-55, expand=True, max_height=60, biome=False)
    _, finds, st = ctx.deepdive(s0, 1 << 20, rect, floor_max=-55, expand=True, max_height=60)
    print("%d spots, %d survivors" % (len(spots), len(surv)), flush=True)
    assert st["funnel"] == st0["funnel"] and st0["biome_checked"] == 0
    want, rejected = biome_filter(surv, 0)
    print("oracle keeps %d, rejects %d" % (len(want), rejected), flush=True)
    assert st["biome_checked"] == len(surv) and st["biome_rejected"] == rejected
    assert finds == want
    assert want and rejected, "both outcomes exercised"
    # the first spots' survivors are the oracle's phase-1 scan at maxHeight 60
    ref = expansion_reference(spots[:4], max_y=60)
    first = {(s[0], s[3], s[4]) for s in spots[:4]}
    assert [h for h in surv if (h[0] & M48, h[3], h[4]) in first] == ref["hits"]


@pytest.mark.parametrize("version", ["1.18", "26.2"])
def test_expansion_versions_match_oracle(ctx, mst, version):
    vflags = whf.version_flags(version)
    f = mst["finds"][3]
    rx, rz = f["cx"] >> 5, f["cz"] >> 5
    spots, finds, st = ctx.deepdive(f["seed"] & M48, 1, (rx, rx + 1, rz, rz + 1), floor_max=-55, expand=True,
                                    max_height=-50, vflags=vflags)
    ref = expansion_reference(spots, vflags=vflags)
    assert st["funnel"] == ref["funnel"] and st["tiers"] == ref["tiers"]
    want, rejected = biome_filter(ref["hits"], vflags)
    assert st["biome_checked"] == len(ref["hits"]) and st["biome_rejected"] == rejected
    assert finds == want


def test_invalid_requests_are_rejected(ctx):
    for kw in (dict(rect=(0, 0, 0, 1)), dict(rect=(0, 1, 0, 1), count=(1 << 48) + 1),
               dict(rect=(0, 1, 0, 1), floor_max=5000), dict(rect=(0, 1, 0, 1), max_height=-500, expand=True),
               dict(rect=(0, 1, 0, 1), vflags=99)):
        args = dict(s_first=0, count=16, floor_max=-55)
        args.update(kw)
        with pytest.raises(whf.WhfError):
            ctx.deepdive(args.pop("s_first"), args.pop("count"), args.pop("rect"), **args)


def test_deepdive_cli_writes_solver_hits_and_finds(tmp_path, mst):
    """bin/deepdive_scan on a find's structure seed: the hits file has the solver's "S,cx,cz,X,1" row and the
    finds file the world seed, chunk, carver seed and room floor."""
    import subprocess
    f = mst["finds"][7]
    rx, rz = f["cx"] >> 5, f["cz"] >> 5
    hits, out = tmp_path / "hits.csv", tmp_path / "finds.csv"
    r = subprocess.run([whf.DEEPDIVE_SCAN, "--s-first", str(f["seed"] & M48), "--count", "1", "--min-region-x", str(rx),
                        "--max-region-x", str(rx + 1), "--min-region-z", str(rz), "--max-region-z", str(rz + 1),
                        "--floor", "-55", "--expand", "--max-height", "-50", "--dump-hits", str(hits), "--out", str(out),
                        "--stats"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    X = O.carver_seed(f["seed"] & M48, f["cx"], f["cz"])
    assert hits.read_text() == "%d,%d,%d,%d,1\n" % (f["seed"] & M48, f["cx"], f["cz"], X)
    lines = out.read_text().split("\n")
    assert lines[0] == "seed,chunkX,chunkZ,layoutX,roomFloor"
    assert "%d,%d,%d,%d,%d" % (f["seed"], f["cx"], f["cz"], X, f["room_floor"]) in lines
