"""CPU oracle of the biome stage against the reference's own outputs and against what the mathematics fixes.

Pins (tests/golden/biome_vectors.json, written by make_biome_golden.py from the reference CPU scanner):
  spline   600 (c, e, w) -> offset/factor/peaks bit patterns, half of them at the spline's own locations
  search   2400 targets -> winner and leaf, per tree, half of them on or next to a leaf's box faces
  point    73 quart points -> shifted coordinates, continentalness, weirdness, erosion, temperature, humidity,
           overworld offset as double bit patterns (a fifth with LARGE_BIOMES)
  check    292 (seed, hut, tree, occupied height) -> six quantized targets, quart position and winner
  heights  25 (seed, hut, tree) -> the winner at all 97 quart heights
Independent of the reference: the tree search returns a leaf at the brute-force minimum distance (branch and
bound over bounding boxes is exact), the depth gradient's closed form, and the version -> tree table.
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
    with open(os.path.join(HERE, "golden", "biome_vectors.json")) as f:
        return json.load(f)


def f32(h):
    return struct.unpack("<f", struct.pack("<I", int(h, 16)))[0]


def fbits(v):
    return "%08x" % struct.unpack("<I", struct.pack("<f", v))[0]


def dbits(v):
    return "%016x" % struct.unpack("<Q", struct.pack("<d", v))[0]


def test_spline_vectors(vec):
    bad = []
    for c, e, w, off, fac, peak in vec["spline"]:
        got = O.terrain_spline(f32(c), f32(e), f32(w))
        if [fbits(v) for v in got] != [off, fac, peak]:
            bad.append((c, e, w, [fbits(v) for v in got], [off, fac, peak]))
    assert not bad, bad[:3]


def test_tree_search_vectors(vec):
    bad = []
    for tree, *rest in vec["search"]:
        t, winner, leaf = rest[:6], rest[6], rest[7]
        if O.tree_leaf(tree, t) != leaf or O.tree_biome(tree, t) != winner:
            bad.append((tree, t, winner, leaf, O.tree_leaf(tree, t)))
    assert not bad, bad[:3]
    assert len({v[0] for v in vec["search"]}) == 8


def test_tree_search_is_nearest_leaf():
    """The walk prunes a subtree only when its box is no nearer than the best leaf so far, so the leaf it returns
    is at the minimum distance over all leaves (ties aside, which leaf wins among equals is the walk's order)."""
    rng = random.Random(7)
    for tree in range(8):
        for _ in range(60):
            t = [rng.randint(-14000, 14000) for _ in range(6)]
            leaf = O.tree_leaf(tree, t)
            dmin, _ = O.tree_min_dist(tree, t)
            assert O.tree_dist(tree, leaf, t) == dmin, (tree, t)


def test_tree_dist_closed_form():
    # inside a leaf's box the distance is 0; one unit past a face it is 1; the root box holds every target
    # within the game's parameter space, so its distance there is 0
    assert O.tree_dist(3, 0, [0, 0, 0, 0, 0, 0]) == 0
    rng = random.Random(3)
    for _ in range(20):
        tree = rng.randrange(8)
        t = [rng.randint(-9000, 9000) for _ in range(6)]
        leaf = O.tree_leaf(tree, t)
        d = O.tree_dist(tree, leaf, t)
        assert d >= 0
        assert O.tree_dist(tree, leaf, t) == O.tree_min_dist(tree, t)[0]


def test_point_vectors(vec):
    bad = []
    cache = {}
    for seed, qx, qz, large, *hexes in vec["point"]:
        key = (seed, large)
        if key not in cache:
            cache[key] = O.Biome(int(seed), O.V_LARGE if large else 0)
        got = [dbits(v) for v in cache[key].point(qx, qz)]
        if got != hexes:
            bad.append((seed, qx, qz, large, [i for i in range(8) if got[i] != hexes[i]]))
    assert not bad, bad[:5]


def test_check_vectors(vec):
    bad = []
    cache = {}
    swamp = 0
    for seed, hx, hz, large, tree, occ, t0, t1, t2, t3, t4, t5, qx, qy, qz, winner in vec["check"]:
        key = (seed, large)
        if key not in cache:
            cache[key] = O.Biome(int(seed), O.V_LARGE if large else 0)
        r = cache[key].check(tree, hx, hz, occ)
        want = dict(t=[t0, t1, t2, t3, t4, t5], qx=qx, qy=qy, qz=qz, winner=winner)
        got = {k: r[k] for k in want}
        if got != want:
            bad.append((seed, hx, hz, tree, occ, got, want))
        assert r["ok"] == (winner == O.SWAMP)
        swamp += winner == O.SWAMP
    assert not bad, bad[:3]


def test_heights_vectors(vec):
    bad = []
    swamp = 0
    for seed, hx, hz, large, tree, winners in vec["heights"]:
        b = O.Biome(int(seed), O.V_LARGE if large else 0)
        got = b.heights(tree, hx, hz)
        if got != winners:
            bad.append((seed, hx, hz, tree))
        swamp += winners.count(O.SWAMP)
        # the height-indexed winners are the check's winners at those heights
        for i in (0, 48, 96):
            assert b.check(tree, hx, hz, (O.BIOME_QY0 + i) * 4)["winner"] == got[i]
    assert not bad, bad[:3]
    assert swamp > 0


def test_depth_gradient_closed_form():
    assert O.depth_gradient(-64.0) == 1.5
    assert O.depth_gradient(320.0) == -1.5
    assert O.depth_gradient(128.0) == 0.0
    assert O.depth_gradient(-1000.0) == 1.5 and O.depth_gradient(1000.0) == -1.5
    for y in range(-64, 321, 4):
        assert abs(O.depth_gradient(float(y)) - (1.5 - 3.0 * (y + 64) / 384.0)) < 1e-15


def test_version_trees():
    # the reference's version table (biome_trees.h:70501-70526) and the rule-set defaults (eval.c:104-106)
    assert O.biome_tree("1.18") == 0 and O.biome_tree("1.18.2") == 0 and O.biome_tree("1.18.x") == 0
    assert O.biome_tree("1.19.2") == 1 and O.biome_tree("1.19.4") == 2
    assert O.biome_tree("1.20") == 3 and O.biome_tree("1.21") == 3 and O.biome_tree("1.21.4") == 4
    assert O.biome_tree("1.21.5") == 5 and O.biome_tree("26.1") == 5 and O.biome_tree("1.21.x~26.1") == 5
    assert O.biome_tree("26.2") == 6 and O.biome_tree("26.2~26.3") == 6 and O.biome_tree("26.3") == 7
    with pytest.raises(ValueError):
        O.biome_tree("1.17")


def test_planted_spline_mistake_is_caught(vec):
    """The spline fixture tells a transposed argument order apart: (c, w, e) instead of (c, e, w)."""
    diff = 0
    for c, e, w, off, fac, peak in vec["spline"][:200]:
        got = O.terrain_spline(f32(c), f32(w), f32(e))
        diff += [fbits(v) for v in got] != [off, fac, peak]
    assert diff > 100
