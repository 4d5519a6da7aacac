"""Parity of the HIP carvers and hut evaluation (whf_post_carver_heights, whf_hut_eval through the C ABI) with the
reference's own outputs and the CPU oracle.

Bar: integers, compared exactly: post- and pre-carver heights, and every field of the hut verdict (orientation,
footprint, biome verdict and occupied height, wet columns, flooded, height sum and average, ok, reject reason).
Inputs: tests/golden/hut_vectors.json (lysh_eval_hut on 76 huts, 60 of 90 columns lowered by the carvers), and
random huts and columns against the oracle.
"""
import json
import os
import random

import pytest

import oracle_ctypes as O
from product import whf

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIELDS = ("dir", "size_x", "size_z", "biome_ok", "winner", "occ_y", "wet", "flooded", "sum_h", "avg_y", "ok",
          "reject")


@pytest.fixture(scope="module")
def ctx():
    print("opening whf context on device 0", flush=True)
    c = whf.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def vec():
    with open(os.path.join(GOLD, "hut_vectors.json")) as f:
        return json.load(f)


def test_post_carver_heights(ctx, vec):
    rows = [(int(s), x, z) for s, x, z, _, _ in vec["post"]]
    print("whf_post_carver_heights: %d columns" % len(rows), flush=True)
    post, pre = ctx.post_carver_heights(rows)
    bad = [(r, (int(a), int(b)), (v[3], v[4])) for r, a, b, v in zip(rows, post, pre, vec["post"]) if (a, b) != (v[3], v[4])]
    assert not bad, bad[:5]


def check(ctx, rows):
    by_tree = {}
    for i, r in enumerate(rows):
        by_tree.setdefault(r[3], []).append(i)
    bad = []
    for tree, idx in sorted(by_tree.items()):
        print("whf_hut_eval tree %d: %d huts" % (tree, len(idx)), flush=True)
        got = ctx.hut_eval([(int(rows[i][0]), rows[i][1] >> 4, rows[i][2] >> 4) for i in idx], tree)
        for i, g in zip(idx, got):
            want = dict(zip(FIELDS, rows[i][4:]))
            if {k: g[k] for k in FIELDS} != want:
                bad.append((rows[i][:4], {k: (g[k], want[k]) for k in FIELDS if g[k] != want[k]}))
    return bad


def test_hut_eval_reference(ctx, vec):
    bad = check(ctx, vec["eval"] + vec["eval2"])
    assert not bad, bad[:3]


def test_hut_eval_random_against_oracle(ctx):
    rng = random.Random(23)
    huts = [(rng.getrandbits(64), rng.randint(-20000, 20000), rng.randint(-20000, 20000)) for _ in range(40)]
    got = ctx.hut_eval(huts, 5)
    bad = []
    for (s, cx, cz), g in zip(huts, got):
        want = O.Terrain(s).hut_eval(5, cx * 16, cz * 16, -50)
        if any(g[k] != want[k] for k in FIELDS):
            bad.append(((s, cx, cz), {k: (g[k], want[k]) for k in FIELDS if g[k] != want[k]}))
    assert not bad, bad[:3]


@pytest.mark.parametrize("version", ["1.18.2", "26.2"])
def test_hut_eval_rule_sets_against_oracle(ctx, version):
    """Huts under the 1.18 and 26.2 rule sets and with LARGE_BIOMES (the terrain and climate noises of the large
    preset), against the oracle."""
    rng = random.Random(37 if version == "26.2" else 41)
    tree = whf.biome_tree(version)
    for large in (0, O.V_LARGE):
        vflags = whf.version_flags(version) | large
        huts = [(rng.getrandbits(64), rng.randint(-20000, 20000), rng.randint(-20000, 20000)) for _ in range(12)]
        print("whf_hut_eval: %s, vflags %d" % (version, vflags), flush=True)
        got = ctx.hut_eval(huts, tree, -50, vflags)
        for (s, cx, cz), g in zip(huts, got):
            want = O.Terrain(s, vflags).hut_eval(tree, cx * 16, cz * 16, -50)
            assert all(g[k] == want[k] for k in FIELDS), ((s, cx, cz), vflags, g, want)


def test_hut_eval_two_rounds(ctx):
    """More huts than one device round holds (512), with repeated seeds and repeated huts: a sample, including the
    rounds' first and last huts, equals the oracle."""
    rng = random.Random(31)
    seeds = [rng.getrandbits(64) for _ in range(40)]
    huts = [(seeds[k % 40], rng.randint(-20000, 20000), rng.randint(-20000, 20000)) for k in range(600)]
    huts[600 - 1] = huts[3]
    print("whf_hut_eval: %d huts" % len(huts), flush=True)
    got = ctx.hut_eval(huts, 3)
    assert got[599] == got[3]
    for i in [0, 1, 510, 511, 512, 513, 598] + rng.sample(range(600), 30):
        s, cx, cz = huts[i]
        want = O.Terrain(s).hut_eval(3, cx * 16, cz * 16, -50)
        assert all(got[i][k] == want[k] for k in FIELDS), (i, huts[i])


def test_post_heights_random_against_oracle(ctx):
    rng = random.Random(29)
    cols = []
    for _ in range(12):
        s, cx, cz = rng.getrandbits(64), rng.randint(-5000, 5000), rng.randint(-5000, 5000)
        cols += [(s, cx * 16 + rng.randint(0, 15), cz * 16 + rng.randint(0, 15)) for _ in range(8)]
    post, pre = ctx.post_carver_heights(cols)
    for (s, x, z), a, b in zip(cols, post, pre):
        t = O.Terrain(s)
        assert (int(a), int(b)) == (t.post_carver_height(x, z), t.height(x, z)[0]), (s, x, z)


def test_invalid(ctx):
    with pytest.raises(whf.WhfError):
        ctx.hut_eval([(1, 0, 0)], 9)
    assert ctx.hut_eval([], 3) == []
