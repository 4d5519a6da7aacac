"""The CPU oracle of the Deep Dive solver (oracle/whf_oracle_solve.c), pinned four ways:

* the reference's own answers: the self-test hut of hunt.py (REFERENCE, hunt.py:83: seed 3159335117276934358 at
  chunk (-1852878, -1380275), layout 254754461452089) and every confirmed deep find of data/finds.tsv whose chunk
  holds a mineshaft start (tests/golden/mstree_vectors.json) come out of their chunk's task;
* planted solutions: for random structure seeds in all three regimes the task of the seed's own hut chunk, with
  the layout its carver seed gives there, must list the seed, for chunks of all three tiers;
* brute force of the search itself: stopped at 24 bits it must keep exactly the states t mod 2^24 that pass the hut
  filter and zero the carver equation's bits below m + 8 (numpy over all 2^24 states);
* the task space: group residues, regions and tier offsets against their definitions (taskspace.h,
  solver_fused.cuh:25-33) computed here independently.
"""
import json
import os

import numpy as np
import pytest

import oracle_ctypes as O

M48 = (1 << 48) - 1
A = 0x5DEECE66D
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REF_HUT = dict(seed=3159335117276934358, cx=-1852878, cz=-1380275, layout=254754461452089)   # hunt.py:83


def valuation(v):
    if v == 0:
        return 99
    k = 0
    while not v & 1:
        v >>= 1
        k += 1
    return k


def tier_of(a, b):
    va, vb = valuation(a), valuation(b)
    if va != vb:
        return 1 if min(va, vb) == 0 else 2
    return 3 if va < 99 else 0


def regime_of(r):
    """The regime digit of a region coordinate (taskspace.h:13-19)."""
    if -16384 <= r <= 16383:
        return 0
    if -32768 <= r <= 32767:
        return 1
    return 2 if -58593 <= r <= 58593 else None


def test_reference_selftest_hut():
    S = REF_HUT["seed"] & M48
    cx, cz, X = REF_HUT["cx"], REF_HUT["cz"], REF_HUT["layout"]
    assert O.carver_seed(S, cx, cz) == X
    assert S in O.solve_task(X, cx, cz)
    rx, rz = cx >> 5, cz >> 5
    assert (regime_of(rx), regime_of(rz)) == (2, 2) and tier_of(cx & 31, cz & 31) == 1
    c20 = O.creg(rx, rz) & 0xFFFFF
    assert O.perm_group(853310) == c20   # the group hunt.py's -58 search reaches it in
    assert (rx, rz) in O.solve_regions(c20, 2, 2)
    assert (cx & 31, cz & 31) in O.solve_offsets(1)


def test_known_deep_finds():
    with open(os.path.join(GOLD, "mstree_vectors.json")) as f:
        finds = [v for v in json.load(f)["finds"] if v["start"]]
    assert len(finds) >= 5
    for v in finds:
        S, cx, cz = v["seed"] & M48, v["cx"], v["cz"]
        if (cx & 31, cz & 31) == (0, 0):
            continue
        X = O.carver_seed(S, cx, cz)
        assert S in O.solve_task(X, cx, cz), v


def planted(n, seed):
    """n (S, cx, cz, X) with the hut of S at (cx, cz) and X its carver seed there, regions over all regimes."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        S = int(rng.integers(0, 1 << 48, dtype=np.uint64))
        rx, rz = (int(v) for v in rng.integers(-58593, 58594, 2))
        cx, cz = O.hut_chunk(S, rx, rz)
        if (cx & 31, cz & 31) == (0, 0):
            continue
        out.append((S, cx, cz, O.carver_seed(S, cx, cz)))
    return out


def test_planted_solutions_all_tiers():
    seen = {1: 0, 2: 0, 3: 0}
    for S, cx, cz, X in planted(150, 11):
        got = O.solve_task(X, cx, cz)
        assert S in got, (S, cx, cz, X)
        for s in got:   # everything listed is a solution
            assert O.carver_seed(s, cx, cz) == X
            assert O.hut_chunk(s, cx >> 5, cz >> 5) == (cx, cz)
        seen[tier_of(cx & 31, cz & 31)] += 1
    assert min(seen.values()) >= 15, seen


def brute_prefixes(X, cx, cz, width=24):
    """All t mod 2^width that pass the first-two-draws filter and zero the carver equation's bits below
    width + m - 16, from the definitions, over every state (numpy, wrapping uint64 arithmetic)."""
    m = min(valuation(cx & 0xFFFFFFFF) if cx else 40, valuation(cz & 0xFFFFFFFF) if cz else 40)
    t = np.arange(1 << width, dtype=np.uint64)
    a2 = (A * A) & M48
    k2 = (0xB * (A + 1)) & M48
    a4 = (a2 * a2) & M48
    k4 = (0xB * (A ** 3 + A * A + A + 1)) & M48
    m48, m32 = np.uint64(M48), np.uint64(0xFFFFFFFF)
    h2 = ((np.uint64(a2) * t + np.uint64(k2)) & m48) >> np.uint64(16)
    h4 = ((np.uint64(a4) * t + np.uint64(k4)) & m48) >> np.uint64(16)
    e = ((t ^ np.uint64(A)) ^ (np.uint64(cx & 0xFFFFFFFF) * h2) ^ (np.uint64(cz & 0xFFFFFFFF) * h4) ^
         np.uint64(X)) & m32
    keep = (e & np.uint64((1 << (width + m - 16)) - 1)) == 0
    creg20 = O.creg(cx >> 5, cz >> 5) & 0xFFFFF
    q = (((t & np.uint64(0xFFFFF)) ^ np.uint64(A & 0xFFFFF)) + np.uint64(creg20)) & np.uint64(0xFFFFF)
    r0 = q ^ np.uint64(A & 0xFFFFF)
    r1 = (r0 * np.uint64(A) + np.uint64(0xB)) & np.uint64(0xFFFFF)
    r2 = (r1 * np.uint64(A) + np.uint64(0xB)) & np.uint64(0xFFFFF)
    keep &= ((r1 >> np.uint64(17)) & np.uint64(7)) == np.uint64((cx & 31) & 7)
    keep &= ((r2 >> np.uint64(17)) & np.uint64(7)) == np.uint64((cz & 31) & 7)
    return sorted(int(v) for v in t[keep])


@pytest.mark.parametrize("case", ["odd", "u2", "eq", "eq4"])
def test_search_keeps_exactly_the_brute_force_prefixes(case):
    # chunks (cx, cz) per tier: valuations (0, 1), (1, 3), (0, 0), (2, 2)
    cx, cz = {"odd": (-1852878 + 1, -1380275 - 1), "u2": (32 * 1234 + 2, 32 * -555 + 8),
              "eq": (32 * 4321 + 5, 32 * -77 + 17), "eq4": (32 * 999 + 4, 32 * -4242 + 12)}[case]
    X = {"odd": 0x123456789AB, "u2": 0xFEDCBA98765, "eq": 0x5DEECE66D00, "eq4": 0x9E3779B97F4}[case]
    want = brute_prefixes(X, cx, cz)
    got = O.solve_prefixes(X, cx, cz, 24)
    assert got == want
    assert len(want) > 100   # the filter leaves a few hundred to a few thousand


def test_group_residues_regions_offsets():
    # permGroup is a bijection of 20-bit values (two odd multiplies and two xor-shifts)
    vals = {O.perm_group(p) for p in range(0, 1 << 20, 97)}
    assert len(vals) == len(range(0, 1 << 20, 97))
    # regions: every (rx, rz) in the regimes with Creg = c20 (mod 2^20), by a dictionary over rz
    c20 = O.perm_group(4242)
    zs = {}
    for rz in list(range(-58593, -32768)) + list(range(32768, 58594)):
        zs[O.creg(0, rz) - O.creg(0, 0) & 0xFFFFF] = rz
    want = []
    for rx in list(range(-58593, -32768)) + list(range(32768, 58594)):
        need = (c20 - O.creg(rx, 0)) & 0xFFFFF
        if need in zs:
            want.append((rx, zs[need]))
    got = O.solve_regions(c20, 2, 2)
    assert got == want and 2300 < len(got) < 2800
    for rx, rz in got[:50]:
        assert O.creg(rx, rz) & 0xFFFFF == c20
    # tiers partition the 575 offsets other than (0, 0)
    tiers = [set(O.solve_offsets(t)) for t in (1, 2, 3)]
    assert sum(len(t) for t in tiers) == 575 and len(tiers[0] | tiers[1] | tiers[2]) == 575
    for t, s in zip((1, 2, 3), tiers):
        assert all(tier_of(a, b) == t for a, b in s)
    assert O.solve_task(1, 32 * 5, 32 * 7) is None   # offset (0, 0): in no tier
