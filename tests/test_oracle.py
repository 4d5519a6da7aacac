"""Pins the CPU oracle (oracle/whf_oracle.c) before it is trusted as the checker of the HIP path.

Each oracle function is checked against at least one of:
  * golden vectors written by the reference CPU scanner itself (tests/golden/lysh_*.json, made by
    tests/golden/make_golden.py from the reference's lysh sources);
  * closed forms and published test vectors (RFC 1321 MD5, SplitMix64, java.util.Random, the reference's
    own known answers in hunt.py / tools/test_hunt.py);
  * invariants the mathematics fixes (permutations, the y = 0 special case, the tier bound).
When the reference build is present (oracle/_ref), the oracle is also compared with it directly.
"""
import json
import os
import struct
import subprocess

import numpy as np
import pytest

import oracle_ctypes as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REF_BENCH = os.path.join(O.REPO, "oracle", "_ref", "lysh_cpu_bench")


@pytest.fixture(scope="module")
def vec():
    with open(os.path.join(GOLD, "lysh_vectors.json")) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def funnels():
    with open(os.path.join(GOLD, "lysh_funnels.json")) as f:
        return json.load(f)


def bits(x):
    return struct.unpack("<Q", struct.pack("<d", x))[0]


# ---------------------------------------------------------------- closed forms

def test_md5_rfc1321_vectors():
    # RFC 1321 appendix A.5 test suite
    assert O.md5(b"") == "d41d8cd98f00b204e9800998ecf8427e"
    assert O.md5(b"a") == "0cc175b9c0f1b6a831c399e269772661"
    assert O.md5(b"abc") == "900150983cd24fb0d6963f7d28e17f72"
    assert O.md5(b"message digest") == "f96b697d7cb7938d525a2f31aaf161d0"
    assert O.md5(b"abcdefghijklmnopqrstuvwxyz") == "c3fcd3d76192e4007dfb496cca67e13b"
    assert O.md5(b"1234567890" * 8) == "57edf4a22be3c955ac49da2e2107b67a"


def test_splitmix64_known_outputs():
    # Vigna's reference splitmix64.c seeded with 0 / 1234567
    assert O.splitmix(0, 0) == 0xE220A8397B1DCDAF
    assert O.splitmix(0, 1) == 0x6E789E6AA1B965F4
    assert O.splitmix(1234567, 0) == 0x599ED017FB08FC85


def test_scrambled_index_matches_hunt_known_answer():
    # tools/test_hunt.py:292 — index 6 of the scrambled stream is 1291329195661654798
    assert O.s64(O.lib().wo_scramble_index(6)) == 1291329195661654798


def test_java_random_known_values():
    # java.util.Random(seed): setSeed scrambles with 0x5DEECE66D (rng.c:189-192)
    def jr(seed):
        st = O.ctypes.c_uint64((seed ^ 0x5DEECE66D) & ((1 << 48) - 1))
        return st
    st = jr(0)
    assert O.s64(O.lib().wo_lcg_next_long(O.ctypes.byref(st))) == -4962768465676381896
    st = jr(42)
    assert O.lib().wo_lcg_next_int_seeded(O.ctypes.byref(st), 10) == 0
    st = jr(42)   # new Random(42).nextLong(): first word is the well-known nextInt() value -1170105035
    assert O.s64(O.lib().wo_lcg_next_long(O.ctypes.byref(st))) == -5025562857975149833
    # nextInt(bound) equals the brute-force definition on many draws (rejection rule of Java's nextInt)
    st = jr(123)
    raw = jr(123)
    for _ in range(2000):
        got = O.lib().wo_lcg_next_int_seeded(O.ctypes.byref(st), 24)
        while True:
            raw.value = (raw.value * 0x5DEECE66D + 0xB) & ((1 << 48) - 1)
            u = raw.value >> 17
            v = u % 24
            if u - v + 23 < 2 ** 31:
                break
        assert got == v


def test_hut_chunk_brute_force():
    # structure.c:4-24 recomputed in Python integers
    M = (1 << 48) - 1
    for seed in [0, 1, -1, 1473884015873122059, 2 ** 62 + 12345]:
        for rx, rz in [(0, 0), (-1, 5), (58593, -58594), (-493, -537)]:
            s = (rx * 341873128712 + rz * 132897987541 + seed + 14357620) & 0xFFFFFFFFFFFFFFFF
            st = (s ^ 0x5DEECE66D) & M
            out = []
            for _ in range(2):
                while True:
                    st = (st * 0x5DEECE66D + 0xB) & M
                    u = st >> 17
                    v = u % 24
                    if u - v + 23 < 2 ** 31:
                        break
                out.append(v)
            assert O.hut_chunk(seed, rx, rz) == (rx * 32 + out[0], rz * 32 + out[1])


def test_known_answer_chunk():
    # hunt.py:86 SCAN_REFERENCE hut at block (-15792, -17168) lies in region (-31, -34)
    assert O.hut_chunk(1473884015873122059, -31, -34) == (-15792 // 16, -17168 // 16)


def test_amplitude_special_values():
    # noise.c:116-122: erosion span 4, continentalness span 8, single octave span 0
    assert O.Noise(1, O.N_EROSION).amplitude == 1.3888888888888888
    assert O.Noise(1, O.N_CONTINENTALNESS).amplitude == 1.4999999999999998
    assert O.Noise(1, O.N_FLOODEDNESS).amplitude == 0.8333333333333333


def test_perm_tables_are_permutations():
    n = O.Noise(0x1234, O.N_CAVE_CHEESE)
    for sub in (0, 1):
        for slot in (0, 1, 2, 3, 4, 5, 7):
            assert sorted(n.perm(sub, slot)) == list(range(256))


def test_y0_is_the_3d_sample_at_zero():
    n = O.Noise(99, O.N_TEMPERATURE)
    rng = np.random.default_rng(0)
    for x, z in rng.uniform(-1e6, 1e6, size=(200, 2)):
        assert bits(n.sample_y0(x, z)) == bits(n.sample(x, 0.0, z))


def test_tiered_erosion_bound_and_exact_gate():
    """No tier fired => the interleaved sum is >= 0.55, and it differs from the exact value by reassociation only;
    a tier that fired at the last threshold means the exact value is below 0.55 (up to reassociation)."""
    n = O.Noise(77, O.N_EROSION)
    rng = np.random.default_rng(1)
    for x, z in rng.uniform(-2e5, 2e5, size=(3000, 2)):
        fired, tier, v = n.tiered(x, z)
        exact = n.sample_y0(x, z)
        if not fired:
            assert v >= 0.55
            assert abs(v - exact) <= 4e-16 * 4
        elif tier == 6:
            assert v < 0.55 and abs(v - exact) <= 1e-15


# ---------------------------------------------------------------- golden vectors from the reference

def test_golden_md5_names(vec):
    for name, digest in vec["md5"]:
        assert O.md5(name.encode()) == digest, name


def test_golden_xoroshiro(vec):
    for row in vec["xr"]:
        s = O.XR()
        O.lib().wo_xr_seed(O.ctypes.byref(s), int(row["seed"]))
        assert [str(O.lib().wo_xr_next(O.ctypes.byref(s))) for _ in range(4)] == row["longs"]
        assert [O.lib().wo_xr_next_int(O.ctypes.byref(s), 256 - 31 * k) for k in range(8)] == row["ints"]
        assert "%016x" % bits(O.lib().wo_xr_next_double(O.ctypes.byref(s))) == row["double"]


def test_golden_placement(vec):
    for seed, rx, rz, cx, cz in vec["lcg"]:
        assert O.hut_chunk(int(seed), rx, rz) == (cx, cz)


def test_golden_sampler_fingerprints(vec):
    for seed, noise, sub, slot, fp in vec["sampler"]:
        if sub < 0:
            assert "%016x" % bits(O.Noise(int(seed), noise).amplitude) == fp
        else:
            assert "%016x" % O.lib().wo_sampler_fingerprint(int(seed), noise, sub, slot) == fp, (seed, noise, sub, slot)


def test_golden_probe_bits(vec):
    for row in vec["probe"]:
        got = ["%016x" % b for b in O.probe(int(row["seed"]), row["x"], row["z"], row["max_y"])]
        assert got == row["v"], (row["seed"], row["x"], row["z"])


def test_golden_region_kinds(vec):
    for row in vec["regions"]:
        vf = O.version_flags(row["version"])
        rect = (row["rx0"], row["rx1"], row["rz0"], row["rz1"])
        kind, tier = O.scan_regions(int(row["seed"]), rect, row["max_y"], vf)
        assert "".join("0123456789A"[k] for k in kind) == row["kinds"]
        assert "".join("-" if t < 0 else str(t) for t in tier) == row["tiers"]


@pytest.mark.parametrize("case", ["c1_v118", "c1_v262", "c2_prefix_v121"])
def test_golden_funnels(funnels, case):
    c = funnels[case]
    r = O.scan(O.seeds_splitmix(c["stream"], c["count"]), c["rect"], c["max_y"], O.version_flags(c["version"]))
    assert r["checks"] == c["checks"]
    assert r["funnel"] == c["funnel"]
    assert r["tiers"] == c["tiers"]
    assert r["reached_ladder"] == c["reached_ladder"]
    assert r["accepted"] == c["accepted"]


def test_golden_known_answer(funnels):
    k = funnels["known_answer"]
    r = O.scan(np.array([k["seed"]], np.int64), k["rect"], k["max_y"], O.version_flags(k["version"]))
    # lysh's --list rows carry the hut's block x, y, z in columns 3-5 (main.c:375); phase 1 must keep each hut
    kept = {(h[3] * 16, h[4] * 16) for h in r["hits"]}
    assert len(k["hits"]) == 1
    for _, _, x, y, z, fl in k["hits"]:
        assert (x, z) in kept
    assert (-15792, -17168) in kept


# ---------------------------------------------------------------- direct comparison with the reference build

@pytest.mark.skipif(not os.path.exists(REF_BENCH), reason="reference build oracle/_ref not present")
def test_oracle_equals_reference_build_random_seeds():
    seed = 987654321
    for ver, rect in [("1.20", (-3, 3, -3, 3)), ("26.3", (-5, 5, -2, 2)), ("1.18.2", (10, 14, -40, -36))]:
        out = subprocess.run([REF_BENCH, "--stream", str(seed), "--count", "2000", "--rx0", str(rect[0]), "--rx1",
                              str(rect[1]), "--rz0", str(rect[2]), "--rz1", str(rect[3]), "--version", ver,
                              "--threads", "4"], check=True, capture_output=True, text=True).stdout
        ref = json.loads(out)
        r = O.scan(O.seeds_splitmix(seed, 2000), rect, -50, O.version_flags(ver), threads=4)
        assert r["funnel"] == ref["funnel"] and r["tiers"] == ref["tiers"]
