"""CPU oracle of the terrain stage against the reference's own outputs and against what the mathematics fixes.

Pins (tests/golden/p2_vectors.json, written by make_p2_golden.py from the reference CPU scanner, 90 columns of 40
seeds, a fifth with LARGE_BIOMES, near and far from the origin, about half of them ocean):
  point 90, blended 180, column 30 x 49, noodle 120 x 4, density 360 (all bit patterns), psl 270, skip 90,
  fluid 135 (level, type), block 540 (state), height 90 (pre-carver height and the state of its top block).
Independent of the reference: the final density is the trilinear blend of the corner columns (it equals them at
the corners), the blended noise's octave sums are bounded, and every surface height sits on a non-air block with
air above it.
"""
import json
import os
import random
import struct

import pytest

import oracle_ctypes as O

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def vec():
    with open(os.path.join(HERE, "golden", "p2_vectors.json")) as f:
        return json.load(f)


_cache = {}


def terrain(seed, large):
    key = (int(seed), large)
    if key not in _cache:
        _cache[key] = O.Terrain(int(seed), O.V_LARGE if large else 0)
    return _cache[key]


def dbits(v):
    return "%016x" % struct.unpack("<Q", struct.pack("<d", v))[0]


def fbits(v):
    return "%08x" % struct.unpack("<I", struct.pack("<f", v))[0]


def test_terrain_point(vec):
    bad = [r for r in vec["point"] if [fbits(v) for v in terrain(r[0], r[1]).point(r[2], r[3])] != r[4:7]]
    assert not bad, bad[:3]


def test_blended_noise(vec):
    bad = [r for r in vec["blended"] if dbits(terrain(r[0], r[1]).blended(r[2], r[3], r[4])) != r[5]]
    assert not bad, bad[:3]


def test_corner_columns(vec):
    bad = []
    for seed, large, cx, cz, col in vec["column"]:
        main, _ = terrain(seed, large).corner_column(cx, cz)
        got = [dbits(v) for v in main]
        if got != col:
            bad.append((seed, cx, cz, [i for i in range(49) if got[i] != col[i]]))
    assert not bad, bad[:3]


def test_noodle(vec):
    bad = []
    for seed, large, x, y, z, *hexes in vec["noodle"]:
        t = terrain(seed, large)
        got = [dbits(t.noodle(w, x, y, z)) for w in range(4)]
        if got != hexes:
            bad.append((seed, x, y, z))
    assert not bad, bad[:3]


def test_final_density(vec):
    bad = [r for r in vec["density"] if dbits(terrain(r[0], r[1]).density(r[2], r[3], r[4])) != r[5]]
    assert not bad, bad[:3]


def test_prelim_surface(vec):
    bad = [r for r in vec["psl"] if terrain(r[0], r[1]).prelim_surface(r[2], r[3]) != r[4]]
    assert not bad, bad[:3]


def test_skip_above(vec):
    bad = [r for r in vec["skip"] if terrain(r[0], r[1]).skip_above(r[2], r[3]) != r[4]]
    assert not bad, bad[:3]


def test_fluid(vec):
    bad = [r for r in vec["fluid"] if terrain(r[0], r[1]).fluid(r[2], r[3], r[4]) != (r[5], r[6])]
    assert not bad, bad[:3]


def test_block_states(vec):
    bad = []
    for seed, large, x, y, z, state in vec["block"]:
        t = terrain(seed, large)
        got = t.substance(x, y, z, t.density(x, y, z))
        if got != state:
            bad.append((seed, x, y, z, got, state))
    assert not bad, bad[:3]


def test_heights(vec):
    bad = []
    for seed, large, x, z, h, top in vec["height"]:
        got = terrain(seed, large).height(x, z)
        if got != (h, top):
            bad.append((seed, x, z, got, (h, top)))
    assert not bad, bad[:3]


def test_density_is_corner_blend():
    """At a corner block the interpolated density is the squeeze of the corner value (min with the noodle)."""
    t = O.Terrain(99)
    main, _ = t.corner_column(10, -7)
    for i in (0, 10, 16, 20, 30):
        v = main[i]
        a = max(-1.0, min(1.0, v * 0.64))
        squeezed = a / 2.0 - a * a * a / 24.0
        assert t.density(40, -64 + 8 * i, -28) <= squeezed + 0.0


def test_surface_is_on_a_block():
    rng = random.Random(4)
    t = O.Terrain(12345)
    for _ in range(6):
        x, z = rng.randint(-2000, 2000), rng.randint(-2000, 2000)
        h, st = t.height(x, z)
        assert st != O.AIR
        assert t.substance(x, h - 1, z, t.density(x, h - 1, z)) == st
        if h < 320:
            assert t.substance(x, h, z, t.density(x, h, z)) == O.AIR


# Seeds whose blended-noise random (createRandom("minecraft:terrain") of the world's deriver, interp_noise.c:19-48)
# rejects a bounded draw of Xoroshiro128++'s nextInt while building sampler k < 39: the samplers after it start one
# or more engine steps later than 259 k. Used by the GPU parity tests of the terrain stage.
BLENDED_REJECTION_SEEDS = [(5219354412081042641, 2), (347959725458519193, 13), (5518476781235771321, 18),
                           (3913600556879318232, 36), (7453945117803196337, 38)]


def first_rejecting_sampler(seed):
    """The first of the 40 blended-noise samplers whose permutation draws reject, or None: the game's Xoroshiro128++
    (seed upgrade, deriver, MD5 of the name) and Java 17's bounded nextInt, restated from their published
    definitions."""
    import hashlib
    m64 = (1 << 64) - 1

    def mix64(z):
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m64
        return z ^ (z >> 31)

    def nxt(s):
        lo, hi = s
        out = ((((lo + hi) & m64) << 17 | ((lo + hi) & m64) >> 47) + lo) & m64
        h = hi ^ lo
        return out, ((((lo << 49) | (lo >> 15)) & m64) ^ h ^ ((h << 21) & m64), ((h << 28) | (h >> 36)) & m64)

    d = hashlib.md5(b"minecraft:terrain").digest()
    t_hi, t_lo = int.from_bytes(d[:8], "big"), int.from_bytes(d[8:], "big")
    l = (seed & m64) ^ 0x6A09E667F3BCC909
    s = (mix64(l), mix64((l - 0x61C8864680B583EB) & m64))
    a, s = nxt(s)
    b, s = nxt(s)
    s = (t_hi ^ a, t_lo ^ b)
    for k in range(40):
        for _ in range(3):
            _, s = nxt(s)
        for i in range(256):
            bound = 256 - i
            lim = (1 << 32) % bound
            out, s = nxt(s)
            rejected = False
            while ((out & 0xFFFFFFFF) * bound) & 0xFFFFFFFF < lim:
                rejected = True
                out, s = nxt(s)
            if rejected:
                return k
    return None


def test_blended_rejection_seeds():
    for seed, k in BLENDED_REJECTION_SEEDS:
        assert first_rejecting_sampler(seed) == k
    assert first_rejecting_sampler(1) is None
