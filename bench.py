#!/usr/bin/env python3
"""Benchmark: the GPU Seed Check at scale on one MI355X, beside the reference CPU scanner on the host.

Workload (one step): 2^26 world seeds from SplitMix64(seed=2), hut regions x, z in [-8, 7] (256 per seed,
1.718e10 hut-region checks), version 1.20 rules, phase-1 maxHeight -50, through the plain FP64 kernel
chain (bit-exact with the reference CPU scanner). The seeds are generated on the device, so the inputs
are resident before the timed region starts.

Prints one JSON line: throughput (hut-region checks/s), the roofline of the dominant kernel (measured
with HIP events on the library's stream during a profiled step after the timed steps), and the CPU
baseline: the reference scanner (lysh, built by oracle/Makefile from the reference's own sources) on a
time-bounded prefix of the same seed stream, one seed per host thread.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--seeds N] [--cpu-seconds S]
"""
import argparse
import json
import os
import subprocess
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

import __graft_entry__  # noqa: E402

METRIC = "hut-region checks/sec, FP32-gate vs FP64, bit-exact finds; p50/p99 latency"
# MI355X peak VALU lane-operations per second: 256 CUs x 4 SIMD x 16 lanes x 2.4 GHz (32-bit integer or FP64
# add/mul issue; the same figure as the 78.6 TFLOP/s FP64 vector peak counted without the FMA factor 2).
PEAK_VALU_OPS = 256 * 4 * 16 * 2.4e9
# Algorithmic cost of one Perlin sampler build: 259 Perlin draws + 12 derivation draws of Xoroshiro128++,
# each 21 32-bit lane operations (18 for the 64-bit state update and output, 3 for the bounded draw).
OPS_PER_SAMPLER = 271 * 21


def cpu_baseline(n_seeds_max, seconds):
    exe = os.path.join(REPO, "oracle", "_ref", "lysh_cpu_bench")
    if not os.path.exists(exe):
        return None
    threads = min(16, os.cpu_count() or 1)
    out = subprocess.run([exe, "--stream", "2", "--count", str(n_seeds_max), "--rx0", "-8", "--rx1", "8",
                          "--rz0", "-8", "--rz1", "8", "--version", "1.20", "--max-y", "-50", "--threads",
                          str(threads), "--time-limit", str(seconds)], check=True, capture_output=True,
                         text=True).stdout
    r = json.loads(out)
    return dict(value=round(r["checks_per_s"], 1), unit="checks/s", cores=threads, kind="reference",
                sample="first %d seeds of the same SplitMix64(2) stream x 256 regions (time-bounded %.0f s), "
                       "reference lysh_scan_rect per seed, one seed per host thread" % (r["seeds"], seconds),
                funnel=r["funnel"], tiers=r["tiers"], seeds=r["seeds"])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--seeds", type=int, default=1 << 26)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each GPU would scan its own seed "
              "range independently); run with --gpus 1", file=sys.stderr)
        return 2

    whf = __graft_entry__.load_whf()
    rect = (-8, 8, -8, 8)
    vflags = whf.version_flags("1.20")
    kw = dict(splitmix=(2, args.seeds), max_height=-50, vflags=vflags, capacity=1 << 20)
    with whf.Context(0) as ctx:
        for _ in range(args.warmup):
            ctx.scan(rect, **kw)
        # whf_scan returns after synchronizing the library's stream, so the wall clock between calls brackets
        # exactly the device work of the K steps
        t0 = time.perf_counter()
        last = None
        for _ in range(args.steps):
            hits, last = ctx.scan(rect, **kw)
        t1 = time.perf_counter()
    secs = (t1 - t0) / args.steps
    checks = last["checks"]

    # profiled step: per-kernel-family device time from HIP events recorded on the library's stream
    with whf.Context(0, profile=True) as ctx:
        _, prof = ctx.scan(rect, **kw)
    fam = prof["kernel_ms"]
    dominant = max(fam, key=fam.get)
    n_batches = -(-args.seeds // (1 << 17))
    built = prof["samplers_built"]
    set_of = {"build_erosion": 0, "build_climate": 1, "build_ladder": 2}
    if dominant in set_of:
        ops = built[set_of[dominant]] * OPS_PER_SAMPLER
        achieved = ops / (fam[dominant] / 1e3)
        roofline = dict(bound="valu", achieved=round(achieved / 1e12, 3), peak=round(PEAK_VALU_OPS / 1e12, 2),
                        unit="Tops/s", frac=round(achieved / PEAK_VALU_OPS, 4), traffic=None,
                        kernel="k_build (%s)" % dominant, launch_ms=round(fam[dominant] / n_batches, 4),
                        ops_per_launch=ops // n_batches)
    else:
        roofline = dict(bound="valu", achieved=None, peak=round(PEAK_VALU_OPS / 1e12, 2), unit="Tops/s", frac=None,
                        traffic=None, kernel=dominant, launch_ms=round(fam[dominant] / n_batches, 4))

    cpu = None if args.no_cpu else cpu_baseline(min(args.seeds, 1 << 20), args.cpu_seconds)
    line = {
        "metric": METRIC, "value": round(checks / secs, 1), "unit": "checks/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(secs * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "Seed Check at scale: %d SplitMix64(2) seeds x regions [-8,7]^2 (256 per seed), "
                               "version 1.20, phase-1 maxHeight -50, plain FP64 kernel chain" % args.seeds,
                   "seeds": args.seeds, "regions_per_seed": 256, "version": "1.20", "precision": "fp64"},
        "roofline": roofline, "cpu_baseline": cpu,
        "funnel": last["funnel"], "tiers": last["tiers"], "accepted": last["accepted"],
        "kernel_ms": {k: round(v, 3) for k, v in fam.items()}, "samplers_built": built,
        "device_ms": round(last["device_ms"], 3),
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
