"""Parity of the HIP biome stage (whf_biome through the C ABI) with the reference's own outputs and the CPU oracle.

Bar: the answers are integers (quantized climate targets, quart positions, biome ids), so everything is exact:
the six targets, the quart position and the winner at the query's height, the winner at all 97 quart heights
and the swamp bit mask built from them.
Inputs: tests/golden/biome_vectors.json (the reference CPU scanner's biome check), phase-1 survivors that the
HIP Seed Check finds, random hut chunks near and far from the origin, all eight trees, both noise presets.
"""
import json
import os
import random

import numpy as np
import pytest

import oracle_ctypes as O
from product import whf

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def ctx():
    print("opening whf context on device 0", flush=True)
    c = whf.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def vec():
    with open(os.path.join(GOLD, "biome_vectors.json")) as f:
        return json.load(f)


def run_grouped(ctx, rows):
    """whf_biome over rows (seed, hut_x, hut_z, occ_y, tree, large), one call per (tree, preset)."""
    out = [None] * len(rows)
    win = [None] * len(rows)
    groups = {}
    for i, r in enumerate(rows):
        groups.setdefault((r[4], r[5]), []).append(i)
    for (tree, large), idx in sorted(groups.items()):
        print("whf_biome tree %d large %d: %d queries" % (tree, large, len(idx)), flush=True)
        res, w = ctx.biome([rows[i][:4] for i in idx], tree, O.V_LARGE if large else 0, winners=True)
        for k, i in enumerate(idx):
            out[i] = res[k]
            win[i] = list(w[k])
    return out, win


def test_check_vectors(ctx, vec):
    rows = [(int(s), hx, hz, occ, tree, large) for s, hx, hz, large, tree, occ, *_ in vec["check"]]
    res, _ = run_grouped(ctx, rows)
    bad = []
    for r, (s, hx, hz, large, tree, occ, t0, t1, t2, t3, t4, t5, qx, qy, qz, winner) in zip(res, vec["check"]):
        got = (r["target"], r["qx"], r["qy"], r["qz"], r["winner"], r["ok"])
        want = ([t0, t1, t2, t3, t4, t5], qx, qy, qz, winner, int(winner == whf.BIOME_SWAMP))
        if got != want:
            bad.append((s, hx, hz, tree, occ, got, want))
    assert not bad, bad[:3]


def test_heights_vectors(ctx, vec):
    rows = [(int(s), hx, hz, 0, tree, large) for s, hx, hz, large, tree, _ in vec["heights"]]
    res, win = run_grouped(ctx, rows)
    for r, w, (s, hx, hz, large, tree, winners) in zip(res, win, vec["heights"]):
        assert w == winners, (s, hx, hz, tree)
        mask = sum(1 << i for i, b in enumerate(winners) if b == whf.BIOME_SWAMP)
        assert r["swamp_qy"] == mask


def oracle_rows(rows):
    cache = {}
    want = []
    for s, hx, hz, occ, tree, large in rows:
        key = (s, large)
        if key not in cache:
            cache[key] = O.Biome(s, O.V_LARGE if large else 0)
        b = cache[key]
        want.append((b.check(tree, hx, hz, occ), b.heights(tree, hx, hz)))
    return want


def compare(rows, res, win, want):
    bad = []
    for row, r, w, (c, h) in zip(rows, res, win, want):
        got = (r["target"], r["qx"], r["qy"], r["qz"], r["winner"], r["ok"], w)
        exp = (c["t"], c["qx"], c["qy"], c["qz"], c["winner"], c["ok"], h)
        if got != exp:
            bad.append((row, got[:6], exp[:6], w == h))
        assert r["swamp_qy"] == sum(1 << i for i, b in enumerate(h) if b == whf.BIOME_SWAMP)
    assert not bad, bad[:3]


def test_random_chunks_match_oracle(ctx):
    rng = random.Random(11)
    rows = []
    for k in range(600):
        seed = rng.getrandbits(64)
        far = k % 3 == 0
        lim = 1874999 if far else 4000
        rows.append((seed, 16 * rng.randint(-lim, lim), 16 * rng.randint(-lim, lim), rng.randint(-70, 330),
                     rng.randrange(8), int(k % 7 == 0)))
    want = oracle_rows(rows)
    res, win = run_grouped(ctx, rows)
    compare(rows, res, win, want)


def test_seed_check_survivors(ctx):
    """The stage's real input: phase-1 survivors of the HIP Seed Check, with several heights per survivor."""
    print("scan for survivors", flush=True)
    # phase-1 maxHeight 60 lets about 1e-5 of the regions through (at -50, about 2e-7)
    hits, _ = ctx.scan((-16, 16, -16, 16), splitmix=(1, 8192), max_height=60, vflags=0, capacity=1 << 14)
    assert len(hits) >= 10
    rows = []
    for h in hits[:400]:
        seed, cx, cz = h[0], h[3], h[4]
        for occ in (-60, 62, 63, 70):
            rows.append((seed & 0xFFFFFFFFFFFFFFFF, 16 * cx, 16 * cz, occ, 3, 0))
    want = oracle_rows(rows)
    res, win = run_grouped(ctx, rows)
    compare(rows, res, win, want)
    assert any(x["ok"] for x in res), "no swamp among phase-1 survivors"


def test_batches_and_repeated_seeds(ctx):
    """More queries than one device round holds, most of them sharing a few seeds."""
    rng = random.Random(5)
    seeds = [rng.getrandbits(64) for _ in range(7)]
    rows = [(seeds[k % 7], 16 * rng.randint(-3000, 3000), 16 * rng.randint(-3000, 3000), rng.randint(-65, 319), 5, 0)
            for k in range(9000)]
    res = ctx.biome([r[:4] for r in rows], 5)
    sample = rng.sample(range(len(rows)), 300) + [0, 8191, 8192, 8999]
    want = oracle_rows([rows[i] for i in sample])
    for i, (c, h) in zip(sample, want):
        r = res[i]
        assert (r["target"], r["qx"], r["qy"], r["qz"], r["winner"]) == (c["t"], c["qx"], c["qy"], c["qz"],
                                                                          c["winner"]), rows[i]
        assert r["swamp_qy"] == sum(1 << j for j, b in enumerate(h) if b == whf.BIOME_SWAMP)


def test_invalid_arguments(ctx):
    with pytest.raises(whf.WhfError):
        ctx.biome([(1, 0, 0, 64)], 8)
    with pytest.raises(whf.WhfError):
        ctx.biome([(1, 0, 0, 64)], -1)
    assert ctx.biome([], 3) == []


@pytest.mark.parametrize("tree,vflags", [(3, 0), (7, whf.V_LARGE)])
def test_occ_only_mode_equals_the_full_answer(ctx, tree, vflags):
    """WHF_BIO_OCC_ONLY (one tree walk per query, for maps) gives every field of the answer at occ_y that the full
    mode gives (swamp_qy is then 0), over 3,000 random queries at given heights and at the pre-carver surface."""
    rng = random.Random(11 + tree)
    qs = [(rng.getrandbits(64) - (1 << 63), rng.randint(-20000, 20000) * 16, rng.randint(-20000, 20000) * 16,
           rng.choice([rng.randint(-64, 300), whf.OCC_SURFACE]) if i % 7 else whf.OCC_SURFACE) for i in range(3000)]
    full = ctx.biome(qs, tree, vflags)
    one = ctx.biome(qs, tree, vflags | whf.BIO_OCC_ONLY)
    for f, o in zip(full, one):
        assert o["swamp_qy"] == 0
        assert {k: v for k, v in o.items() if k != "swamp_qy"} == {k: v for k, v in f.items() if k != "swamp_qy"}
    assert sum(o["ok"] for o in one) >= 1
    with pytest.raises(whf.WhfError):
        ctx.biome(qs[:4], tree, whf.BIO_OCC_ONLY, winners=True)
