"""Parity of the HIP terrain stage (whf_surface_heights, and whf_biome with WHF_OCC_SURFACE, through the C ABI) with
the reference's own outputs and the CPU oracle.

Bar: the outputs are integers (surface height, top block state, biome id), compared exactly. The FP64 values behind
them follow the reference operation for operation except the pillar thickness cube, which the device rounds once
where the reference's C library pow can be one ulp off; the heights here are all equal regardless.
Inputs: tests/golden/p2_vectors.json (the reference's pre-carver heights, 90 columns, half of them sea), random
columns near and far from the origin with both noise presets, the chunk middles of phase-1 survivors.
"""
import json
import os
import random

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
    with open(os.path.join(GOLD, "p2_vectors.json")) as f:
        return json.load(f)


def heights_grouped(ctx, rows):
    """rows (seed, x, z, large) -> [(height, state)] with one call per preset."""
    out = [None] * len(rows)
    for large in (0, 1):
        idx = [i for i, r in enumerate(rows) if r[3] == large]
        if not idx:
            continue
        print("whf_surface_heights large %d: %d columns" % (large, len(idx)), flush=True)
        h, st = ctx.surface_heights([rows[i][:3] for i in idx], O.V_LARGE if large else 0)
        for k, i in enumerate(idx):
            out[i] = (int(h[k]), int(st[k]))
    return out


def test_reference_heights(ctx, vec):
    rows = [(int(s), x, z, large) for s, large, x, z, _, _ in vec["height"]]
    got = heights_grouped(ctx, rows)
    bad = [(r, g, (v[4], v[5])) for r, g, v in zip(rows, got, vec["height"]) if g != (v[4], v[5])]
    assert not bad, bad[:5]


def test_random_columns_match_oracle(ctx):
    rng = random.Random(17)
    rows = []
    for k in range(160):
        lim = 29_000_000 if k % 4 == 0 else 20_000
        rows.append((rng.getrandbits(64), rng.randint(-lim, lim), rng.randint(-lim, lim), int(k % 6 == 0)))
    got = heights_grouped(ctx, rows)
    bad = []
    for r, g in zip(rows, got):
        want = O.Terrain(r[0], O.V_LARGE if r[3] else 0).height(r[1], r[2])
        if g != want:
            bad.append((r, g, want))
    assert not bad, bad[:5]


def test_biome_at_the_surface(ctx):
    """whf_biome with the height left to the library equals lysh's check with its pre-carver height callback."""
    print("scan for survivors", flush=True)
    hits, _ = ctx.scan((-16, 16, -16, 16), splitmix=(1, 4096), max_height=60, vflags=0, capacity=1 << 14)
    rng = random.Random(3)
    qs = [(h[0] & 0xFFFFFFFFFFFFFFFF, 16 * h[3], 16 * h[4]) for h in hits[:40]]
    qs += [(rng.getrandbits(64), 16 * rng.randint(-3000, 3000), 16 * rng.randint(-3000, 3000)) for _ in range(40)]
    tree = whf.biome_tree("1.21")
    res = ctx.biome([(s, hx, hz, whf.OCC_SURFACE) for s, hx, hz in qs], tree)
    bad = []
    for (s, hx, hz), r in zip(qs, res):
        h, _ = O.Terrain(s).height((hx >> 4) * 16 + 8, (hz >> 4) * 16 + 8)
        want = O.Biome(s).check(tree, hx, hz, h - 1)
        if (r["occ_y"], r["target"], r["winner"]) != (h - 1, want["t"], want["winner"]):
            bad.append(((s, hx, hz), r["occ_y"], h - 1, r["winner"], want["winner"]))
    assert not bad, bad[:3]
    assert any(r["ok"] for r in res[:len(hits[:40])]), "no swamp among the survivors"


def test_many_columns_few_seeds(ctx):
    """More columns than one device round holds, on a few seeds; spot-checked against the oracle."""
    rng = random.Random(8)
    seeds = [rng.getrandbits(64) for _ in range(3)]
    rows = [(seeds[k % 3], rng.randint(-5000, 5000), rng.randint(-5000, 5000)) for k in range(1100)]
    h, st = ctx.surface_heights(rows)
    for i in [0, 1, 1023, 1024, 1099] + rng.sample(range(1100), 20):
        want = O.Terrain(rows[i][0]).height(rows[i][1], rows[i][2])
        assert (int(h[i]), int(st[i])) == want, rows[i]


def test_invalid_and_empty(ctx):
    h, st = ctx.surface_heights([])
    assert len(h) == 0 and len(st) == 0


def test_blended_rejection_seeds(ctx):
    """Seeds whose blended-noise draws reject a bounded draw before the last sampler (the device's per-sampler jump
    ahead is then off, and it builds the samplers again in sequence): heights equal the oracle's."""
    from test_oracle_p2 import BLENDED_REJECTION_SEEDS
    rng = random.Random(23)
    rows = [(seed, rng.randint(-3000, 3000), rng.randint(-3000, 3000)) for seed, _ in BLENDED_REJECTION_SEEDS
            for _ in range(6)]
    # mixed with seeds that take the parallel build, in one round
    rows += [(rng.getrandbits(64), rng.randint(-3000, 3000), rng.randint(-3000, 3000)) for _ in range(10)]
    h, st = ctx.surface_heights(rows)
    for i, r in enumerate(rows):
        assert (int(h[i]), int(st[i])) == O.Terrain(r[0]).height(r[1], r[2]), r
