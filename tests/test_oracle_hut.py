"""CPU oracle of the carvers and the hut evaluation against the reference's own outputs.

Pins (tests/golden/hut_vectors.json, written by make_hut_golden.py from the reference CPU scanner):
  eval   58 lysh_eval_hut results (phase-1 survivors, the known-answer hut, random chunks; trees 0, 3, 6):
         orientation, footprint, biome verdict and occupied height, wet columns, flooded, height sum, average,
         ok and the reject reason
  eval2  18 more on chunks where the carvers lower columns
  post   90 post- and pre-carver column heights, 60 of them lowered by the carvers
Independent of the reference: carving never raises a column, and the orientation is uniform over 0..3.
"""
import collections
import json
import os
import random

import pytest

import oracle_ctypes as O

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def vec():
    with open(os.path.join(HERE, "golden", "hut_vectors.json")) as f:
        return json.load(f)


_cache = {}


def terrain(seed):
    if seed not in _cache:
        _cache[seed] = O.Terrain(seed)
    return _cache[seed]


def check_rows(rows):
    bad = []
    for seed, hx, hz, tree, d, sx, sz, bok, win, occ, wet, fl, sum_h, avg, ok, rej in rows:
        got = terrain(int(seed)).hut_eval(tree, hx, hz, -50)
        want = dict(dir=d, size_x=sx, size_z=sz, biome_ok=bok, winner=win, occ_y=occ, wet=wet, flooded=fl,
                    sum_h=sum_h, avg_y=avg, ok=ok, reject=rej)
        g = {k: got[k] for k in want}
        if g != want:
            bad.append((seed, hx, hz, {k: (g[k], want[k]) for k in want if g[k] != want[k]}))
    return bad


def test_hut_eval(vec):
    bad = check_rows(vec["eval"])
    assert not bad, bad[:3]


def test_hut_eval_on_carved_chunks(vec):
    bad = check_rows(vec["eval2"])
    assert not bad, bad[:3]


def test_post_carver_heights(vec):
    bad = []
    for seed, x, z, post, pre in vec["post"]:
        t = terrain(int(seed))
        got = (t.post_carver_height(x, z), t.height(x, z)[0])
        if got != (post, pre):
            bad.append((seed, x, z, got, (post, pre)))
    assert not bad, bad[:3]
    assert sum(1 for r in vec["post"] if r[3] < r[4]) >= 50


def test_carving_never_raises(vec):
    for seed, x, z, post, pre in vec["post"]:
        assert post <= pre


def test_orientation_is_uniform():
    rng = random.Random(1)
    c = collections.Counter(O.hut_orientation(rng.getrandbits(64), 16 * rng.randint(-9999, 9999),
                                              16 * rng.randint(-9999, 9999)) for _ in range(4000))
    assert set(c) == {0, 1, 2, 3} and min(c.values()) > 850
