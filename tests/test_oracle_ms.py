"""CPU checks of the Deep Dive oracle (oracle/whf_oracle_ms.c): what pins the start test, the carver seed and the
mineshaft room floor before the GPU path is compared with it.

Pins:
  - the reference's own tool: mstree (engine/tools/mstree/mstree.cpp) vectors in tests/golden/mstree_vectors.json,
    written by tests/golden/make_mstree_golden.py: start flags of random chunks, and for start chunks the room floor,
    the piece count and every piece box (SHA-256 of mstree's lines; three small layouts in full);
  - the reference's data/finds.tsv: its 50 confirmed deep huts with coordinates (in the same fixture);
  - closed forms: java.util.Random written out here (setSeed scramble, next(bits), nextLong, nextDouble) for the
    carver seed and the start test; the exact 0.004 threshold as a rational;
  - invariants of the move below sea level: room floor = r - minY - 12 with 0 <= r < max(1, 116 - span), and the
    early cut at -12 - T leaves every "floor <= T" decision unchanged.
"""
import hashlib
import json
import os
from fractions import Fraction

import pytest

import oracle_ctypes as O

HERE = os.path.dirname(os.path.abspath(__file__))
V = json.load(open(os.path.join(HERE, "golden", "mstree_vectors.json")))
M48 = (1 << 48) - 1
MUL = 0x5DEECE66D


# java.util.Random, written out (Java SE API specification of Random)
class JavaRandom:
    def __init__(self, seed):
        self.s = (seed ^ MUL) & M48

    def next(self, bits):
        self.s = (self.s * MUL + 0xB) & M48
        v = self.s >> (48 - bits)
        return v - (1 << bits) if bits == 32 and v >= 1 << 31 else v

    def next_long(self):
        return ((self.next(32) << 32) + self.next(32)) & ((1 << 64) - 1)

    def next_double(self):
        return Fraction((self.next(26) << 27) + self.next(27), 1 << 53)


def signed(v):
    v &= (1 << 64) - 1
    return v - (1 << 64) if v >= 1 << 63 else v


def test_carver_seed_closed_form():
    import random
    rng = random.Random(5)
    for _ in range(500):
        s48 = rng.getrandbits(48)
        cx, cz = rng.randint(-1875000, 1874999), rng.randint(-1875000, 1874999)
        r = JavaRandom(s48)
        a, b = signed(r.next_long()), signed(r.next_long())
        want = ((cx * a) ^ (cz * b) ^ s48) & M48
        assert O.carver_seed(s48, cx, cz) == want


def test_start_test_closed_form_and_threshold():
    import random
    rng = random.Random(6)
    thr = Fraction(0.004)   # the double nearest 0.004, exactly
    assert thr * (1 << 53) == Fraction(1152921504606847, 32)
    n_start = 0
    for _ in range(20000):
        X = rng.getrandbits(48)
        d = JavaRandom(X).next_double()
        assert O.start_double(X) == float(d)
        assert O.is_start(X) == (d < thr)
        n_start += d < thr
    assert 40 <= n_start <= 125   # 0.004 of 20000 = 80


def test_start_flags_match_mstree():
    for v in V["random"] + V["starts"]:
        X = O.carver_seed(v["seed"] & M48, v["cx"], v["cz"])
        assert O.is_start(X) == v["start"], v
    assert sum(v["start"] for v in V["random"]) < 10   # random chunks: starts are rare


def test_layouts_match_mstree():
    for v in V["starts"] + [f for f in V["finds"] if f["start"]]:
        X = O.carver_seed(v["seed"] & M48, v["cx"], v["cz"])
        st, floor, lines, r = O.ms_layout(X, v["cx"], v["cz"])
        assert st == 0
        assert floor == v["room_floor"], v
        assert len(lines) == v["pieces"], v
        assert hashlib.sha256("\n".join(lines).encode()).hexdigest() == v["sha256"], v
    for v in V["layouts"]:
        X = O.carver_seed(v["seed"] & M48, v["cx"], v["cz"])
        st, floor, lines, r = O.ms_layout(X, v["cx"], v["cz"])
        assert lines == v["lines"]
        assert floor == v["room_floor"]


def test_finds_are_deep_mineshaft_rooms_under_huts():
    """Every confirmed deep hut of data/finds.tsv sits in its region's hut chunk, which is a mineshaft start with a
    room floor of -60..-62, below the -55 line of the Deep Dive."""
    assert len(V["finds"]) == 50
    for f in V["finds"]:
        rx, rz = f["cx"] >> 5, f["cz"] >> 5
        assert O.hut_chunk(f["seed"], rx, rz) == (f["cx"], f["cz"])
        assert f["start"]
        X = O.carver_seed(f["seed"] & M48, f["cx"], f["cz"])
        st, floor, pieces, r = O.room_floor(X)
        assert st == 0 and floor == f["room_floor"] <= -60
        assert f["y"] > floor and f["depth"] <= -55


def test_finds_pass_phase1_at_the_hut():
    """The same finds pass lysh's phase-1 chain at maxHeight -50: the expansion's check keeps them."""
    for f in V["finds"][:12]:
        rx, rz = f["cx"] >> 5, f["cz"] >> 5
        kind, tier = O.scan_regions(f["seed"], (rx, rx + 1, rz, rz + 1), max_y=-50)
        assert int(kind[0]) == 10, f


@pytest.mark.parametrize("T", [-50, -55, -58, -61])
def test_room_floor_move_and_cut(T):
    import random
    rng = random.Random(T)
    n = 0
    while n < 1500:
        X = rng.getrandbits(48)
        if not O.is_start(X):
            continue
        n += 1
        st, floor, pieces, r = O.room_floor(X)
        assert st == 0
        # floor = 50 + l - maxY with l = span - 63 + r: r = floor + minY + 12
        rr = floor + r.min_y + 12
        assert 0 <= rr < max(1, 116 - r.span_y)
        assert r.min_y <= 50 and r.max_y >= 54
        cut = -12 - T
        st2, floor2, pieces2, r2 = O.room_floor(X, cut)
        keep_full = floor <= T
        keep_cut = st2 == 0 and floor2 <= T
        assert keep_full == keep_cut
        if st2 == 0:
            assert (floor2, pieces2) == (floor, pieces)
        else:
            assert st2 == 1 and r.min_y < cut


def test_layout_piece_boxes_are_in_the_room_frame():
    import random
    rng = random.Random(11)
    n = 0
    while n < 200:
        s48 = rng.getrandbits(48)
        cx, cz = rng.randint(-1000, 1000), rng.randint(-1000, 1000)
        X = O.carver_seed(s48, cx, cz)
        if not O.is_start(X):
            continue
        n += 1
        st, floor, lines, r = O.ms_layout(X, cx, cz)
        room = lines[0].split()
        assert room[0] == "R" and int(room[1]) == 16 * cx + 2 and int(room[3]) == 16 * cz + 2
        assert int(room[2]) == floor
        st0, floor0, pieces0, r0 = O.room_floor(X)
        assert (st0, floor0, pieces0) == (st, floor, len(lines))   # translation invariance: the walk is relative


def test_sweep_counts_and_spots():
    s0 = O.splitmix(3, 0) & M48
    n = 1 << 13
    out = O.deepdive_sweep(s0, n, (-2, 2, -2, 2), floor_max=-55, threads=4)
    assert out["structure_seeds"] == n and out["tests"] == 16 * n
    starts = spots = 0
    want = []
    for i in range(n):
        s48 = (s0 + i) & M48
        for rz in range(-2, 2):
            for rx in range(-2, 2):
                cx, cz = O.hut_chunk(s48, rx, rz)
                X = O.carver_seed(s48, cx, cz)
                if not O.is_start(X):
                    continue
                starts += 1
                st, floor, pieces, r = O.room_floor(X)
                if st == 0 and floor <= -55:
                    want.append((s48, rx, rz, cx, cz, X, floor, pieces))
    assert out["starts"] == starts
    assert out["spots"] == sorted(want)
    assert 0.0025 < starts / out["tests"] < 0.0055


@pytest.mark.skipif(not os.path.exists(os.path.join(HERE, "..", "oracle", "_ref", "mstree_cpu_bench")),
                    reason="reference Deep Dive harness not built (needs the reference sources)")
def test_reference_generator_sweep_agrees():
    """The reference's own generator (mineshaft_tree_iter.h with abortBelow, in oracle/_ref/mstree_cpu_bench) and
    the oracle count the same starts, walk outcomes and spots."""
    import subprocess
    exe = os.path.join(HERE, "..", "oracle", "_ref", "mstree_cpu_bench")
    s0 = O.splitmix(3, 0) & M48
    for floor_max, rect in ((-55, (-2, 2, -2, 2)), (-50, (-3, 1, 0, 2))):
        out = json.loads(subprocess.run([exe, "--s-first", str(s0), "--count", "16384", "--rx0", str(rect[0]), "--rx1",
                                         str(rect[1]), "--rz0", str(rect[2]), "--rz1", str(rect[3]), "--floor",
                                         str(floor_max), "--threads", "2"], check=True, capture_output=True,
                                        text=True).stdout)
        ref = O.deepdive_sweep(s0, 16384, rect, floor_max=floor_max, cut=-12 - floor_max, threads=4)
        assert (out["starts"], out["walks"], out["spots"]) == (ref["starts"], ref["status"], len(ref["spots"]))
