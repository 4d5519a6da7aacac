"""Parity of the HIP Deep Dive solver (whf_solve, whf_solve_chunks and deepdive_scan --solve-only --solve5, through
the C ABI) with the CPU oracle, with the reference's own answers and with planted solutions.

Bar: integer work, so the hit sets are equal exactly. The oracle (oracle/whf_oracle_solve.c) is pinned in
tests/test_oracle_solve.py; its tasks here are chunks of real groups of the reference's -58 plan (tier "odd"), of
the "u2" and "eq" tiers, and planted hut chunks.
"""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle_ctypes as O
from product import whf

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
M48 = (1 << 48) - 1
with open(os.path.join(GOLD, "solver_inputs.json")) as _f:
    INPUTS = json.load(_f)
HUT = INPUTS["selftest_hut"]
REF_GROUP = 853310   # the group of the self-test hut's region in a regime-22 bundle


@pytest.fixture(scope="module")
def ctx():
    print("opening whf context on device 0", flush=True)
    c = whf.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def group_of():
    """The inverse of permGroup over all 2^20 indices (numpy restatement of taskspace.h:24)."""
    p = np.arange(1 << 20, dtype=np.uint64)
    x = (p * np.uint64(0x9E3779B1)) & np.uint64(0xFFFFF)
    x ^= x >> np.uint64(9)
    x = (x * np.uint64(0x85EBCA6B)) & np.uint64(0xFFFFF)
    x ^= x >> np.uint64(7)
    inv = np.zeros(1 << 20, np.int64)
    inv[x.astype(np.int64)] = p.astype(np.int64)
    return inv


def tier_of(a, b):
    def v(n):
        return 99 if n == 0 else (n & -n).bit_length() - 1
    va, vb = v(a), v(b)
    if va != vb:
        return 1 if min(va, vb) == 0 else 2
    return 3 if va < 99 else 0


def regime_of(r):
    if -16384 <= r <= 16383:
        return 0
    if -32768 <= r <= 32767:
        return 1
    return 2


def test_selftest_hut_group(ctx):
    X, S = HUT["layout"], HUT["seed"] & M48
    print("solving group %d of layout %d" % (REF_GROUP, X), flush=True)
    hits, st = ctx.solve([(X, 2, 2, 1)], REF_GROUP, 1)
    assert (0, S, HUT["cx"], HUT["cz"], X) in hits
    c20 = O.creg(HUT["cx"] >> 5, HUT["cz"] >> 5) & 0xFFFFF
    n_reg = len(O.solve_regions(c20, 2, 2))
    assert st["groups"] == 1 and st["regions"] == n_reg and st["tasks"] == n_reg * 288
    assert st["hits"] == len(hits) > 100
    for _, s, cx, cz, x in hits:
        assert O.carver_seed(s, cx, cz) == X and O.hut_chunk(s, cx >> 5, cz >> 5) == (cx, cz)
    print("%d hits, %d candidates, %.1f ms" % (len(hits), st["candidates"], st["device_ms"]), flush=True)


def group_chunks(c20, regime_x, regime_z, tier, n_regions=None):
    regs = O.solve_regions(c20, regime_x, regime_z)[:n_regions]
    offs = O.solve_offsets(tier)
    return [(32 * rx + a, 32 * rz + b) for rx, rz in regs for a, b in offs]


@pytest.mark.parametrize("tier,layout", [(1, 231564844655014), (2, 156186219368766), (3, 270840896143627)])
def test_chunks_match_oracle(ctx, tier, layout):
    c20 = O.perm_group(777 + tier)
    chunks = group_chunks(c20, 2, 2, tier, n_regions=12 if tier == 1 else 24)
    print("tier %d: %d chunks" % (tier, len(chunks)), flush=True)
    hits, st = ctx.solve_chunks(layout, chunks)
    ref = O.solve_tasks(layout, chunks)
    want = sorted((i, s, chunks[i][0], chunks[i][1]) for i, seeds in enumerate(ref) for s in seeds)
    assert hits == want
    assert st["tasks"] == len(chunks)
    print("%d hits over %d chunks" % (len(hits), len(chunks)), flush=True)


def test_planted_chunks(ctx):
    rng = np.random.default_rng(5)
    seen = {1: 0, 2: 0, 3: 0}
    n = 0
    while n < 90:
        S = int(rng.integers(0, 1 << 48, dtype=np.uint64))
        rx, rz = (int(v) for v in rng.integers(-58593, 58594, 2))
        cx, cz = O.hut_chunk(S, rx, rz)
        if (cx & 31, cz & 31) == (0, 0):
            continue
        X = O.carver_seed(S, cx, cz)
        hits, _ = ctx.solve_chunks(X, [(cx, cz)])
        assert (0, S, cx, cz) in hits, (S, cx, cz, X)
        assert [h[1] for h in hits] == O.solve_task(X, cx, cz)
        seen[tier_of(cx & 31, cz & 31)] += 1
        n += 1
    assert min(seen.values()) >= 8, seen


@pytest.mark.timeout(300)
def test_group_equals_its_chunks(ctx):
    """A whole group of the -58 plan: the group solve (regions and pools on the device) equals the chunk solve of
    every (region, offset) of the group, whose regions come from the oracle; a sample of chunks also against the
    oracle itself."""
    X = int(INPUTS["58"]["plan"].split()[0])
    g = 31337
    c20 = O.perm_group(g)
    hits, st = ctx.solve([(X, 2, 2, 1)], g, 1)
    chunks = group_chunks(c20, 2, 2, 1)
    print("group %d: %d chunks, %d hits" % (g, len(chunks), len(hits)), flush=True)
    assert st["tasks"] == len(chunks)
    ch, cst = ctx.solve_chunks(X, chunks, capacity=1 << 16)
    assert sorted((s, cx, cz) for _, s, cx, cz, _ in hits) == sorted((s, cx, cz) for _, s, cx, cz in ch)
    assert cst["candidates"] == st["candidates"]
    rng = np.random.default_rng(1)
    hit_chunks = {(h[2], h[3]) for h in hits[:40]}
    pick = sorted(set(int(i) for i in rng.integers(0, len(chunks), 400)) |
                  {i for i, c in enumerate(chunks) if c in hit_chunks})
    sample = [chunks[i] for i in pick]
    ref = O.solve_tasks(X, sample)
    want = sorted((s, c[0], c[1]) for c, seeds in zip(sample, ref) for s in seeds)
    in_sample = set(sample)
    got = sorted((s, cx, cz) for _, s, cx, cz, _ in hits if (cx, cz) in in_sample)
    assert got == want and len(want) >= 40


def test_known_deep_finds_from_their_groups(ctx, group_of):
    """Every confirmed deep find of the reference's data/finds.tsv whose chunk holds a mineshaft start comes out of
    the group solve of its own layout, regimes and tier."""
    with open(os.path.join(GOLD, "mstree_vectors.json")) as f:
        finds = [v for v in json.load(f)["finds"] if v["start"]]
    n = 0
    for v in finds:
        S, cx, cz = v["seed"] & M48, v["cx"], v["cz"]
        rx, rz = cx >> 5, cz >> 5
        t = tier_of(cx & 31, cz & 31)
        if t == 0:
            continue
        X = O.carver_seed(S, cx, cz)
        g = int(group_of[O.creg(rx, rz) & 0xFFFFF])
        print("find seed=%d chunk=(%d,%d) group %d tier %d" % (v["seed"], cx, cz, g, t), flush=True)
        hits, _ = ctx.solve([(X, regime_of(rx), regime_of(rz), t)], g, 1)
        assert (0, S, cx, cz, X) in hits, v
        n += 1
    assert n >= 40


def test_bundle_boundaries_and_counts(ctx):
    """Groups across the boundary of two bundles (the -59 plan): each bundle's hits carry its index and layout, and
    a range solved in two calls equals it solved in one."""
    plan = whf.parse_plan(INPUTS["59"]["plan"])
    assert len(plan) == 2
    first = (1 << 20) - 2
    hits, st = ctx.solve(plan, first, 4)
    a, _ = ctx.solve(plan, first, 2)
    b, _ = ctx.solve(plan, first + 2, 2)
    assert hits == sorted(a + b) and st["groups"] == 4
    assert {h[0] for h in hits} == {0, 1}
    for bi, s, cx, cz, x in hits:
        assert x == plan[bi][0]


def test_cli_solve5_drop_in(ctx, tmp_path):
    """hunt.py's solver command line (hunt.py:947-952 with its shipped defaults, hunt.py:76-81) on the deepdive_scan
    binary: the hits file is the group solve's, line for line, and holds the self-test hut."""
    plan = tmp_path / "plan.txt"
    cat = tmp_path / "layout_catalog.txt"
    plan.write_text("%d 22 odd\n" % HUT["layout"])
    cat.write_text("%d %d %d\n" % (HUT["layout"], HUT["air_hi"], HUT["air_lo"]))
    hits_f = tmp_path / "hits.csv"
    cmd = [whf.DEEPDIVE_SCAN, "--catalog", str(cat), "--solve-only", "--solve5", "--plan", str(plan),
           "--fused-blk", "64", "--fused-grid", "0", "--fused-gpb", "128", "--fused-slots", "4", "--fused-threads", "4",
           "--regime", "22", "--first-index", str(REF_GROUP - 1), "--count", "3", "--task-batch", "16384", "--quiet",
           "--dump-hits", str(hits_f), "--out", str(tmp_path / "dummy.csv")]
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "solve:" in r.stdout and "host-check failures 0" in r.stdout
    lines = hits_f.read_text().splitlines()
    want, _ = ctx.solve([(HUT["layout"], 2, 2, 1)], REF_GROUP - 1, 3)
    assert lines == ["%d,%d,%d,%d,1" % (s, cx, cz, x) for _, s, cx, cz, x in want]
    assert "%d,%d,%d,%d,1" % (HUT["seed"] & M48, HUT["cx"], HUT["cz"], HUT["layout"]) in lines
