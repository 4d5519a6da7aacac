"""Parity of the HIP Seed Check (through the C ABI) with the CPU oracle and the reference's golden vectors.

Bit-exact bar: raw phase-1 values are compared as IEEE bit patterns, survivors as sets, and the funnel
(reject kind per region) and erosion tier histograms as exact counts.
"""
import json
import os
import subprocess
import sys

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
    with open(os.path.join(GOLD, "lysh_vectors.json")) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def funnels():
    with open(os.path.join(GOLD, "lysh_funnels.json")) as f:
        return json.load(f)


def same_counts(gpu, ref):
    assert gpu["checks"] == ref["checks"]
    assert gpu["funnel"] == ref["funnel"], (gpu["funnel"], ref["funnel"])
    assert gpu["tiers"] == ref["tiers"], (gpu["tiers"], ref["tiers"])
    assert gpu["reached_ladder"] == ref["reached_ladder"]
    assert gpu["accepted"] == ref["accepted"]


# ---------------------------------------------------------------- raw values

def test_probe_bits_match_reference_golden(ctx, vec):
    rows = vec["probe"]
    for maxh in (-50, -40):
        sel = [r for r in rows if r["max_y"] == maxh]
        seeds = np.array([int(r["seed"]) for r in sel], dtype=np.uint64).view(np.int64)
        got = ctx.probe(seeds, [r["x"] for r in sel], [r["z"] for r in sel], maxh, 0)
        for r, g in zip(sel, got):
            assert ["%016x" % int(v) for v in g] == r["v"], (r["seed"], r["x"], r["z"])


@pytest.mark.parametrize("vflags", [0, 1, 2, 4])
def test_probe_bits_match_oracle_random_positions(ctx, vflags):
    rng = np.random.default_rng(100 + vflags)
    n = 96
    seeds = rng.integers(-2 ** 63, 2 ** 63 - 1, size=n, dtype=np.int64)
    xs = (rng.integers(-30_000_000, 30_000_000, size=n) // 16 * 16).astype(np.int32)
    zs = (rng.integers(-30_000_000, 30_000_000, size=n) // 16 * 16).astype(np.int32)
    xs[:8] = np.arange(8) * 16
    zs[:8] = -np.arange(8) * 16
    got = ctx.probe(seeds, xs, zs, -50, vflags)
    for i in range(n):
        assert [int(v) for v in got[i]] == O.probe(int(seeds[i]), int(xs[i]), int(zs[i]), -50, vflags), i


# ---------------------------------------------------------------- funnels against the reference's own numbers

@pytest.mark.parametrize("case", ["c1_v120", "c1_v118", "c1_v262", "c2_prefix_v120", "c2_prefix_v121"])
def test_funnel_equals_reference_golden(ctx, funnels, case):
    c = funnels[case]
    print("case", case, flush=True)
    hits, st = ctx.scan(tuple(c["rect"]), splitmix=(c["stream"], c["count"]), max_height=c["max_y"],
                        vflags=whf.version_flags(c["version"]))
    same_counts(st, c)


def test_known_answer_survivor(ctx, funnels):
    k = funnels["known_answer"]
    hits, st = ctx.scan(tuple(k["rect"]), seeds=[k["seed"]], max_height=-50, vflags=whf.version_flags(k["version"]))
    ref = O.scan(np.array([k["seed"]], np.int64), k["rect"], -50, O.version_flags(k["version"]))
    same_counts(st, ref)
    assert hits == [tuple(h) for h in ref["hits"]]
    assert (1473884015873122059, -31, -34, -987, -1073) in hits


# ---------------------------------------------------------------- against the oracle: modes, presets, shapes

@pytest.mark.parametrize("vflags", [0, whf.V_118, whf.V_262, whf.V_LARGE, whf.V_SINGLE, whf.V_LARGE | whf.V_262])
def test_presets_match_oracle(ctx, vflags):
    seeds = O.seeds_splitmix(77, 512)
    rect = (-5, 5, -4, 6)
    hits, st = ctx.scan(rect, seeds=seeds, max_height=-50, vflags=vflags)
    ref = O.scan(seeds, rect, -50, vflags)
    same_counts(st, ref)
    assert hits == [tuple(h) for h in ref["hits"]]


@pytest.mark.parametrize("max_h", [-50, -40, -25, 0, 64, 100])
def test_max_heights_match_oracle(ctx, max_h):
    seeds = O.seeds_splitmix(5, 1024)
    rect = (-4, 4, -4, 4)
    hits, st = ctx.scan(rect, seeds=seeds, max_height=max_h, vflags=0)
    ref = O.scan(seeds, rect, max_h, 0)
    same_counts(st, ref)
    assert hits == [tuple(h) for h in ref["hits"]]


def test_seed_modes_agree(ctx):
    stream, n = 0xDEADBEEF, 3000
    a_hits, a = ctx.scan((-2, 2, -2, 2), splitmix=(stream, n, 17))
    b_hits, b = ctx.scan((-2, 2, -2, 2), seeds=O.seeds_splitmix(stream, n, 17))
    assert a_hits == b_hits and a["funnel"] == b["funnel"] and a["tiers"] == b["tiers"]
    # counter form: splitmix64 finalizer of base + i + gamma (the reference's --seed-gen)
    base = 123456789
    c_hits, c = ctx.scan((-2, 2, -2, 2), counter=(base, 500))
    # the SplitMix64 finalizer applied to base + i + gamma, i.e. output 0 of a generator seeded with base + i
    seeds = np.array([O.s64(O.splitmix(base + i, 0)) for i in range(500)], np.int64)
    ref = O.scan(seeds, (-2, 2, -2, 2), -50, 0)
    same_counts(c, ref)


def test_edge_shapes(ctx):
    # empty seed list, one region, a thin row, negative-only and far regions
    _, st = ctx.scan((0, 1, 0, 1), seeds=np.zeros(0, np.int64))
    assert st["checks"] == 0 and sum(st["funnel"]) == 0
    for rect in [(0, 1, 0, 1), (-40, 40, 7, 8), (-9, -3, -12, -10), (58590, 58594, -58594, -58590)]:
        seeds = O.seeds_splitmix(9, 300)
        hits, st = ctx.scan(rect, seeds=seeds)
        ref = O.scan(seeds, rect, -50, 0)
        same_counts(st, ref)
        assert hits == [tuple(h) for h in ref["hits"]]


def test_wide_rect_one_seed_multi_chunk(ctx):
    # 600 x 600 regions: several region chunks inside the erosion kernel
    seed = np.array([1473884015873122059], np.int64)
    rect = (-300, 300, -300, 300)
    hits, st = ctx.scan(rect, seeds=seed)
    ref = O.scan(seed, rect, -50, 0, threads=1)
    same_counts(st, ref)
    assert hits == [tuple(h) for h in ref["hits"]]


def test_invalid_requests(ctx):
    with pytest.raises(whf.WhfError) as e:
        ctx.scan((3, 3, 0, 1), seeds=[1])
    assert e.value.code == whf.EINVAL
    with pytest.raises(whf.WhfError):
        ctx.scan((0, 1, 0, 1), seeds=[1], vflags=99)
    with pytest.raises(whf.WhfError):
        ctx.scan((-70000, 1, 0, 1), seeds=[1])


# ---------------------------------------------------------------- full size: properties

@pytest.mark.timeout(900)
def test_full_c2_properties(ctx, funnels):
    """2^26 SplitMix64(2) seeds x 256 regions (the benchmark workload): counters are internally consistent,
    the 2^14-seed prefix equals the reference's golden funnel, and splitting the seed range changes nothing."""
    rect = (-8, 8, -8, 8)
    n = 1 << 26
    hits, st = ctx.scan(rect, splitmix=(2, n), capacity=1 << 20)
    assert st["checks"] == n * 256
    assert sum(st["funnel"]) == st["checks"]
    assert sum(st["tiers"]) == st["funnel"][0]
    assert st["reached_ladder"] == sum(st["funnel"][3:])
    assert st["accepted"] == st["funnel"][10] == len(hits)
    half = n // 2
    h1, s1 = ctx.scan(rect, splitmix=(2, half), capacity=1 << 20)
    h2, s2 = ctx.scan(rect, splitmix=(2, half, half), capacity=1 << 20)
    assert [a + b for a, b in zip(s1["funnel"], s2["funnel"])] == st["funnel"]
    assert [a + b for a, b in zip(s1["tiers"], s2["tiers"])] == st["tiers"]
    assert sorted(h1 + h2) == hits
    g = funnels["c2_prefix_v120"]
    _, sp = ctx.scan(rect, splitmix=(2, g["count"]))
    same_counts(sp, g)
    # the FP32-gated chain gives the identical answer at full size
    h32, s32 = ctx.scan(rect, splitmix=(2, n), capacity=1 << 20, precision=whf.PREC_FP32_GATE)
    assert h32 == hits
    for k in ("funnel", "tiers", "reached_ladder", "accepted", "checks"):
        assert s32[k] == st[k], k
    # every survivor is a survivor of the oracle too (oracle re-checks each hit's seed over the whole rect)
    for seed in sorted({h[0] for h in hits})[:64]:
        ref = O.scan(np.array([seed], np.int64), rect, -50, 0, threads=1)
        assert [h for h in hits if h[0] == seed] == [tuple(x) for x in ref["hits"]]


# ---------------------------------------------------------------- FP32 gate: same answers as FP64

@pytest.mark.parametrize("case", ["c1_v120", "c1_v262", "c2_prefix_v120"])
def test_fp32_gate_equals_reference_golden(ctx, funnels, case):
    c = funnels[case]
    hits, st = ctx.scan(tuple(c["rect"]), splitmix=(c["stream"], c["count"]), max_height=c["max_y"],
                        vflags=whf.version_flags(c["version"]), precision=whf.PREC_FP32_GATE)
    same_counts(st, c)


@pytest.mark.parametrize("vflags", [0, whf.V_118, whf.V_LARGE, whf.V_SINGLE])
def test_fp32_gate_equals_fp64(ctx, vflags):
    rect = (-20, 20, -20, 20)
    kw = dict(splitmix=(31, 4096), max_height=-50, vflags=vflags)
    h64, s64 = ctx.scan(rect, **kw)
    h32, s32 = ctx.scan(rect, precision=whf.PREC_FP32_GATE, **kw)
    assert h32 == h64
    for k in ("funnel", "tiers", "reached_ladder", "accepted", "checks"):
        assert s32[k] == s64[k], k


def test_fp32_gate_rechecks_near_thresholds(ctx):
    """The known-answer block: FP32 gate output equals the oracle's, and some undecided items took the FP64
    re-evaluation path (it is exercised, not dead code)."""
    seeds = O.seeds_splitmix(12, 2048)
    rect = (-16, 16, -16, 16)
    hits, st = ctx.scan(rect, seeds=seeds, precision=whf.PREC_FP32_GATE)
    ref = O.scan(seeds, rect, -50, 0)
    same_counts(st, ref)
    assert hits == [tuple(h) for h in ref["hits"]]
    assert st["fp64_rechecks"] > 0


def test_fp32_error_within_stated_bound(ctx):
    seeds = O.seeds_splitmix(44, 2048)
    err = ctx.noise_error(seeds, 2048)
    bound = whf.fp32_bound()
    print("per-octave max |fp32 - fp64|:", err, "bound", bound, flush=True)
    assert all(0.0 < e < bound for e in err), (err, bound)
