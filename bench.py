#!/usr/bin/env python3
"""Benchmark: the GPU Seed Check at scale on one MI355X, beside the reference CPU scanner on the host.

Workload (one step): 2^26 world seeds from SplitMix64(seed=2), hut regions x, z in [-8, 7] (256 per seed,
1.718e10 hut-region checks), version 1.20 rules, threshold -52 'or lower': phase 1 at maxHeight -50, then
lysh's hut verdict (biome at the surface, flood test, 63-column post-carver average y <= -52) on every
phase-1 survivor, so a step ends with the (seed, chunk, floor) finds, bit-exact with the reference CPU
scanner. The seeds are generated on the device, so the inputs are resident before the timed region starts.

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
# MI355X published peaks (dense vector): FP32 157.3 TFLOP/s, FP64 78.6 TFLOP/s. The 32-bit VALU lane-operation
# issue rate, 256 CUs x 4 SIMD x 16 lanes x 2.4 GHz = 39.3 T/s, bounds the integer work of the sampler builds (the
# published FP64 figure counts an FMA as 2 FLOP at this same issue rate).
PEAK_FP32 = 157.3e12
PEAK_FP64 = 78.6e12
PEAK_VALU_OPS = 256 * 4 * 16 * 2.4e9
# Algorithmic work of one y = 0 Perlin evaluation (noise.c:292-318 with the octave accumulation noise.c:441-449):
# the corner blend is 51 FLOP (8 gradients x 2, 2 fades x 7 with the y fade precomputed, 7 lerps x 3) and the
# coordinates and accumulation 8 FLOP (2 origin adds, 2 fraction subtractions, weight multiply and add, amplitude
# scale, lacunarity scale). The FP32-gated erosion kernel does the blend in FP32 and the rest in FP64; the plain
# kernel does all 59 in FP64.
FLOP_BLEND_PER_EVAL = 51
FLOP_COORD_PER_EVAL = 8
# One 3D Perlin evaluation: 58 FLOP (8 gradients x 2, 3 fades x 7, 7 lerps x 3) and 12 FLOP of coordinates.
FLOP_3D_PER_EVAL = 58
FLOP_3D_COORD = 12
# Algorithmic integer work of one Perlin sampler build: 259 Xoroshiro128++ draws of the permutation and origin
# (noise.c:61-78), each about 25 32-bit lane operations (64-bit state update and output on 32-bit halves, the
# bounded draw's multiply-high), plus the 256 byte swaps.
OPS_PER_SAMPLER = 259 * 25
# The same figure as round numbers per Xoroshiro step, kept for the secondary lane-op fields.
OPS_PER_Y0_EVAL_LANE = 4 + 2 + 2 + 4 + 8 * 5 + 2 * 7 + 7 * 3 + 3
KERNEL_OF = {"seeds": "k_seeds", "build_erosion": "k_build (set E)", "erosion_gate": "k_erosion<4, false>",
             "compact": "k_compact", "build_climate": "k_build (set C)", "gate2": "k_gate2a + k_gate2b",
             "build_ladder": "k_build (set L)", "gate3": "k_gate3"}


def note(msg):
    """A progress line on stderr after each finished step or stage (the JSON result goes to stdout)."""
    print("[bench %.1fs] %s" % (time.perf_counter() - T0, msg), file=sys.stderr, flush=True)


T0 = time.perf_counter()


def erosion_evals(st):
    """y = 0 Perlin evaluations of the tiered erosion gate in one scan, from the scan's own counters: tier t fires
    after t + 2 evaluations, and a region that passes the gate takes all 8."""
    return sum((t + 2) * n for t, n in enumerate(st["tiers"])) + 8 * (st["checks"] - st["funnel"][0])


def gate_evals(st, max_height, single_biome=False):
    """Perlin evaluations of gates 2 and 3 in one scan, as lysh's phase-1 chain makes them (phase1.c:115-200), from
    the scan's funnel, per kernel family: {"gate2": (y0, 3d), "gate3": (y0, 3d)}. Octaves per double noise (noise_table.h, non-zero
    amplitudes, two samplers each): temperature 4, ridge 6, continentalness 18 y = 0 evaluations; Entrance 18,
    Entrance2 12, Cheese 16, floodedness 2 3D evaluations (caves.c:99-131). A Cheese runs only where the Entrance or
    Entrance2 beside it is >= 0; the funnel shows that only for the regions the pair rejects, so the other Cheese
    calls are not counted and the figures are lower bounds. A region rejected inside a ladder is counted at its first
    level."""
    f = st["funnel"]
    r_temp = st["checks"] - f[0]
    r_ridge = r_temp - f[1]
    r_e50 = r_ridge - f[2]
    r_e60 = r_e50 - f[3]
    r_maxh = r_e60 - f[4]
    # gate 2 (k_gate2a, k_gate2b): temperature, ridge, Entrance@50 and @60; gate 3 (k_gate3): cave@maxHeight, the
    # ladders, continentalness and floodedness
    g2 = (4 * r_temp + 6 * r_ridge, 18 * r_e50 + 18 * r_e60)
    ladder_floor = max(-50, min(-40, max_height))
    neg_levels = sum(1 for y in range(0, ladder_floor - 1, -10) if max_height < y)
    flood_levels = len(range(max_height, 61, 10))
    r_neg = r_maxh - f[5]
    r_pos = r_neg - f[6]
    r_cont = r_pos - f[7]
    r_flood = r_cont - (0 if single_biome else f[8])
    g3_3d = 12 * r_maxh + 16 * f[5] + (12 * neg_levels * (r_neg - f[6]) + (12 + 16) * f[6] if neg_levels else 0) + \
        18 * 4 * (r_pos - f[7]) + (18 + 16) * f[7] + 2 * flood_levels * (r_flood - f[9]) + 2 * f[9]
    g3_y0 = 0 if single_biome else 18 * r_cont
    return {"gate2": g2, "gate3": (g3_y0, g3_3d)}


def family_work(st, fp32_gate, max_height=-50):
    """Algorithmic work per kernel family of one scan: (fp32 FLOP, fp64 FLOP, int32 lane-ops). Gates 2 and 3 run in
    FP64: 59 FLOP per y = 0 evaluation, 70 per 3D evaluation."""
    built = st["samplers_built"]
    ev = erosion_evals(st)
    ero = (ev * FLOP_BLEND_PER_EVAL, ev * FLOP_COORD_PER_EVAL, 0) if fp32_gate else \
        (0, ev * (FLOP_BLEND_PER_EVAL + FLOP_COORD_PER_EVAL), 0)
    out = {"build_erosion": (0, 0, built[0] * OPS_PER_SAMPLER), "build_climate": (0, 0, built[1] * OPS_PER_SAMPLER),
           "build_ladder": (0, 0, built[2] * OPS_PER_SAMPLER), "erosion_gate": ero}
    for fam, (y0, d3) in gate_evals(st, max_height).items():
        out[fam] = (0, y0 * (FLOP_BLEND_PER_EVAL + FLOP_COORD_PER_EVAL) + d3 * (FLOP_3D_PER_EVAL + FLOP_3D_COORD), 0)
    return out


def roofline_of(work, seconds):
    """Fractions of the published peaks: FP32 and FP64 FLOP/s against 157.3 and 78.6 TFLOP/s, integer lane-ops
    against the 39.3 T/s issue rate. The kernel's fraction of the vector unit is their sum (the three kinds of work
    share one issue port)."""
    f32, f64, i32 = (w / seconds for w in work)
    fr = dict(fp32=f32 / PEAK_FP32, fp64=f64 / PEAK_FP64, int32=i32 / PEAK_VALU_OPS)
    return dict(fp32_tflops=round(f32 / 1e12, 3), fp64_tflops=round(f64 / 1e12, 3), int32_tops=round(i32 / 1e12, 3),
                frac_fp32=round(fr["fp32"], 4), frac_fp64=round(fr["fp64"], 4), frac_int32=round(fr["int32"], 4),
                frac=round(sum(fr.values()), 4))


def traffic_from_profiles(kernel):
    """HBM bytes per launch of a kernel (or of "a + b", kernels launched together as one family) from the committed
    rocprofv3 PMC summary (FETCH_SIZE and WRITE_SIZE scaled by the copy-kernel calibration), or None when it was not
    measured."""
    path = os.path.join(REPO, "profiles", "traffic_c2.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    parts = [t.get(k) for k in kernel.split(" + ")]
    return None if None in parts else sum(parts)


def host_cpus():
    """The host's CPUs as this process sees them: nproc (affinity), os.cpu_count(), the cgroup CPU quota, the CPU
    model, and the thread count the CPU baseline uses (the smallest of affinity and quota: every core this process
    may run on)."""
    affinity = len(os.sched_getaffinity(0))
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, period = f.read().split()
        if q != "max":
            quota = int(q) / int(period)
    except (OSError, ValueError):
        pass
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    threads = affinity if quota is None else max(1, min(affinity, int(quota)))
    return dict(nproc=affinity, cpu_count=os.cpu_count(), cgroup_quota_cpus=quota, model=model, threads=threads)


def cpu_baseline(n_seeds_max, seconds, host):
    exe = os.path.join(REPO, "oracle", "_ref", "lysh_cpu_bench")
    if not os.path.exists(exe):
        return None
    threads = host["threads"]
    out = subprocess.run([exe, "--stream", "2", "--count", str(n_seeds_max), "--rx0", "-8", "--rx1", "8",
                          "--rz0", "-8", "--rz1", "8", "--version", "1.20", "--max-y", "-50", "--threads",
                          str(threads), "--time-limit", str(seconds), "--phase2-max-y", str(THRESHOLD), "--list", "1"],
                         check=True, capture_output=True, text=True).stdout
    r = json.loads(out)
    return dict(value=round(r["checks_per_s"], 1), unit="checks/s", cores=threads, kind="reference",
                sample="first %d seeds of the same SplitMix64(2) stream x 256 regions (time-bounded %.0f s), "
                       "reference lysh_scan_rect per seed with its phase-2 hook (lysh_eval_hut, max-y %d) on the "
                       "survivors, one seed per host thread" % (r["seeds"], seconds, THRESHOLD),
                host=dict(host, cflags=r["cflags"], compiler="gcc " + r["compiler"]),
                funnel=r["funnel"], tiers=r["tiers"], seeds=r["seeds"], checks=r["checks"], accepted=r["accepted"],
                reached_ladder=r["reached_ladder"], phase2_evaluated=r["phase2_evaluated"],
                phase2_accepted=r["phase2_accepted"], phase2_rejected=r["phase2_rejected"],
                finds=[tuple(f) for f in r["finds"]],
                phase2_ms_per_hut=round(1e3 * r["phase2_seconds"] / r["phase2_evaluated"], 3)
                if r["phase2_evaluated"] else None)


def autotune(seeds):
    """Runs the autotuner (low-witch-finder_amd/tune.py) for versions 1.18, 1.20 and 1.21 into a file of its own and
    points whf_open at it (WHF_TUNED), before any timed context opens; returns what it chose and measured."""
    import importlib.util
    import tempfile
    spec = importlib.util.spec_from_file_location("whf_tune", os.path.join(REPO, "low-witch-finder_amd", "tune.py"))
    tune = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(tune)
    fd, path = tempfile.mkstemp(prefix="whf_tuned_", suffix=".json")
    os.close(fd)
    os.unlink(path)
    t0 = time.perf_counter()
    device, entry = tune.main(["--out", path, "--seeds", str(seeds), "--versions", "1.18,1.20,1.21"])
    secs = time.perf_counter() - t0
    os.environ["WHF_TUNED"] = path
    per = entry["measured"]["per_version"]
    return dict(seconds=round(secs, 2), device=device, build=entry["build"],
                chosen={v: p["chosen"] for v, p in per.items()},
                default_ms={v: p["default_ms"] for v, p in per.items()},
                tuned_ms={v: p["tuned_ms"] for v, p in per.items()},
                chunk_checks=entry.get("chunk_checks"), workload=entry["measured"]["workload"],
                serving_capacity=entry["measured"].get("serving_capacity_checks_per_s"))


def cpu_parity(whf, cpu, precisions):
    """The GPU scan of exactly the seeds the CPU baseline scanned (its stream prefix), with the same threshold and
    version, compared with the reference scanner's own counters and finds: checks, the reject funnel, the erosion tier
    histogram, phase-1 survivors and cells reaching the cave ladder, the phase-2 verdict counters and the list of finds
    (seed, chunk x, chunk z, floor). One comparison per GPU precision."""
    rect = (-8, 8, -8, 8)
    out = dict(seeds=cpu["seeds"], compared=["checks", "funnel", "tiers", "accepted", "reached_ladder",
                                             "phase2 counters", "finds (seed, cx, cz, floor)"])
    want_finds = sorted(cpu["finds"])
    want_p2 = [cpu["phase2_evaluated"], cpu["phase2_accepted"]] + list(cpu["phase2_rejected"])
    equal = True
    with whf.Context(0) as ctx:
        for prec in precisions:
            hits, st = ctx.scan(rect, splitmix=(2, cpu["seeds"]), max_height=-50, vflags=whf.version_flags("1.20"),
                                capacity=1 << 20, precision=prec,
                                phase2=(whf.P2_OR_LOWER, THRESHOLD, whf.biome_tree("1.20")))
            got_finds = sorted((h[0], h[3], h[4], h[5]) for h in hits)
            diff = [k for k, a, b in (("checks", st["checks"], cpu["checks"]), ("funnel", st["funnel"], cpu["funnel"]),
                                      ("tiers", st["tiers"], cpu["tiers"]), ("accepted", st["accepted"], cpu["accepted"]),
                                      ("reached_ladder", st["reached_ladder"], cpu["reached_ladder"]),
                                      ("phase2", st["phase2"], want_p2), ("finds", got_finds, want_finds)) if a != b]
            name = "fp64" if prec == whf.PREC_FP64 else "fp32_gate"
            out[name] = dict(equal=not diff, differs=diff, finds=len(got_finds), phase2=st["phase2"])
            equal = equal and not diff
    out["equal"] = equal
    out["finds"] = [list(f) for f in want_finds]
    return out


def cpu_product_mode(n_seeds, threads):
    """hunt.py's CPU mode: one `lysh scan --list --no-stats` process per seed (hunt.py:1390-1392), `threads` processes
    at a time, over the first n_seeds seeds of the SplitMix64(2) stream x 256 regions, threshold -52."""
    from concurrent.futures import ThreadPoolExecutor
    exe = os.path.join(REPO, "oracle", "_ref", "lysh")
    if not os.path.exists(exe):
        return None
    seeds = [splitmix_at(2, i) for i in range(n_seeds)]

    def one(seed):
        return subprocess.run([exe, "scan", "--seed", str(seed), "--rx0", "-8", "--rx1", "8", "--rz0", "-8", "--rz1",
                               "8", "--threads", "1", "--max-y", "-50", "--phase2-max-y", str(THRESHOLD), "--version",
                               "1.20", "--list", "--no-stats"], check=True, capture_output=True, text=True).stdout

    t0 = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        outs = list(ex.map(one, seeds))
    secs = time.perf_counter() - t0
    hits = sum(o.count("\nhit ") + o.startswith("hit ") for o in outs)
    return dict(value=round(n_seeds * 256 / secs, 1), unit="checks/s", cores=threads, kind="reference",
                sample="one `lysh scan --list --no-stats --threads 1` process per seed, %d processes at a time, over "
                       "the first %d seeds of the same stream x 256 regions, phase-2 max-y %d" % (threads, n_seeds,
                                                                                            THRESHOLD),
                seconds=round(secs, 3), ms_per_seed_per_process=round(1e3 * secs * threads / n_seeds, 3), hits=hits)


THRESHOLD = -52      # the Seed Check's floor threshold, 'or lower' (phase 1 runs at maxHeight -50)
DD_RECT = (-2, 2, -2, 2)
DD_FLOOR = -55
DD_TOTAL = 1 << 32   # structure seeds of the full Deep Dive workload


def splitmix_at(stream, i):
    """Output i (from 0) of sequential SplitMix64 seeded with `stream` (Vigna's splitmix64.c), as a signed 64-bit
    world seed."""
    m = (1 << 64) - 1
    z = (stream + (i + 1) * 0x9E3779B97F4A7C15) & m
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    z ^= z >> 31
    return z - (1 << 64) if z >> 63 else z


def splitmix_seeds(stream, count):
    """Outputs 0 .. count-1 of sequential SplitMix64 seeded with `stream`, as signed 64-bit world seeds (numpy)."""
    import numpy as np
    with np.errstate(over="ignore"):
        z = np.uint64(stream) + (np.arange(1, count + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z ^= z >> np.uint64(31)
    return z.view(np.int64)


def splitmix_first(stream):
    """First output of sequential SplitMix64 seeded with `stream`, unsigned."""
    return splitmix_at(stream, 0) & ((1 << 64) - 1)


def deep_dive_cpu(s0, seconds):
    exe = os.path.join(REPO, "oracle", "_ref", "mstree_cpu_bench")
    if not os.path.exists(exe):
        return None
    threads = host_cpus()["threads"]
    out = subprocess.run([exe, "--s-first", str(s0), "--count", str(1 << 26), "--rx0", str(DD_RECT[0]), "--rx1",
                          str(DD_RECT[1]), "--rz0", str(DD_RECT[2]), "--rz1", str(DD_RECT[3]), "--floor", str(DD_FLOOR),
                          "--threads", str(threads), "--time-limit", str(seconds)], check=True, capture_output=True,
                         text=True).stdout
    r = json.loads(out)
    return dict(value=round(r["seeds_per_s"], 1), unit="structure seeds/s (start tests + room-floor walks)",
                cores=threads, kind="reference",
                sample="first %d structure seeds of the same stream, regions [-2,1]^2, the reference's mineshaft "
                       "generator (mineshaft_tree_iter.h) with its abortBelow cut, time-bounded %.0f s"
                       % (r["structure_seeds"], seconds), starts=r["starts"], spots=r["spots"])


# The Deep Dive solver's workload: the -58 search's plan (data/hunt/58/plan.txt: one bundle, layout 231564844655014,
# regime 22, tier "odd"), groups from 2^17 on (hunt.py skips the first eighth of the first bundle, hunt.py:886-891).
SOLVER_PLAN = [(231564844655014, 2, 2, 1)]
SOLVER_FIRST = 1 << 17
# Algorithmic integer work of the solver, per candidate (a low-20-bit state that passes the pool filters): the lift
# of its 28 - m remaining bits, each a bit test, two conditional 32-bit adds and the xor of three running values
# (6 lane-ops), and per full state the carver equation's bits 32..47 (two 48-bit and two 32-bit LCG products, two
# high and two low 32-bit multiplies, xors: about 30 lane-ops).
SOLVER_OPS_PER_LEVEL = 6
SOLVER_OPS_PER_LEAF = 30


def solver_cpu(seconds, threads):
    """The oracle's solver (oracle/whf_oracle_solve.c) on host threads over chunks of the workload's first group,
    time-bounded: tasks per second."""
    import ctypes
    so = os.path.join(REPO, "oracle", "_build", "libwhf_oracle.so")
    if not os.path.exists(so):
        return None
    L = ctypes.CDLL(so)
    L.wo_solve_regions.restype = ctypes.c_int
    L.wo_solve_regions.argtypes = [ctypes.c_uint32, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int32),
                                   ctypes.POINTER(ctypes.c_int32), ctypes.c_int]
    L.wo_solve_perm_group.restype = ctypes.c_uint32
    L.wo_solve_perm_group.argtypes = [ctypes.c_uint32]
    L.wo_solve_offsets.restype = ctypes.c_int
    L.wo_solve_offsets.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_uint8), ctypes.POINTER(ctypes.c_uint8)]
    L.wo_solve_tasks.restype = ctypes.c_int64
    L.wo_solve_tasks.argtypes = [ctypes.c_uint64, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                 ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_uint64),
                                 ctypes.POINTER(ctypes.c_int)]
    X, rgx, rgz, tier = SOLVER_PLAN[0]
    rx = (ctypes.c_int32 * 60000)()
    rz = (ctypes.c_int32 * 60000)()
    nr = L.wo_solve_regions(L.wo_solve_perm_group(SOLVER_FIRST & 0xFFFFF), rgx, rgz, rx, rz, 60000)
    ox = (ctypes.c_uint8 * 576)()
    oz = (ctypes.c_uint8 * 576)()
    nj = L.wo_solve_offsets(tier, ox, oz)
    done, hits = 0, 0
    t0 = time.perf_counter()
    batch = 64 * threads
    while time.perf_counter() - t0 < seconds and done < nr * nj:
        n = min(batch, nr * nj - done)
        cx = (ctypes.c_int32 * n)(*[32 * rx[(done + i) // nj] + ox[(done + i) % nj] for i in range(n)])
        cz = (ctypes.c_int32 * n)(*[32 * rz[(done + i) // nj] + oz[(done + i) % nj] for i in range(n)])
        seeds = (ctypes.c_uint64 * (n * 16))()
        counts = (ctypes.c_int * n)()
        hits += L.wo_solve_tasks(X, cx, cz, n, 16, threads, seeds, counts)
        done += n
    secs = time.perf_counter() - t0
    return dict(value=round(done / secs, 1), unit="tasks/s", cores=threads, kind="port",
                sample="the first %d (region, offset) tasks of group %d of the -58 plan, time-bounded %.0f s: the "
                       "oracle's solver (every low-20-bit state, then a bit-serial lift), one task per host thread"
                       % (done, SOLVER_FIRST, seconds), tasks=done, hits=int(hits))


def solver(whf, groups, cpu_seconds, host):
    """The Deep Dive solver (whf_solve) over `groups` groups of the -58 plan, as one call after a one-group warm-up."""
    with whf.Context(0) as ctx:
        ctx.solve(SOLVER_PLAN, SOLVER_FIRST, 1)
        t0 = time.perf_counter()
        hits, st = ctx.solve(SOLVER_PLAN, SOLVER_FIRST, groups)
        secs = time.perf_counter() - t0
    note("solver: %d groups in %.3f s, %d hits" % (groups, secs, len(hits)))
    work = st["candidates"] * 28 * SOLVER_OPS_PER_LEVEL + st["leaves"] * SOLVER_OPS_PER_LEAF
    dev_s = st["device_ms"] / 1e3
    cpu = None if cpu_seconds <= 0 else solver_cpu(cpu_seconds, host["threads"])
    out = dict(workload="the -58 plan's bundle (layout 231564844655014, regime 22, tier odd), groups [%d, %d): "
                        "hunt.py's solver stage (hut_lookup_v4 --solve-only --solve5)" % (SOLVER_FIRST,
                                                                                        SOLVER_FIRST + groups),
               groups=groups, regions=st["regions"], tasks=st["tasks"], candidates=st["candidates"],
               leaves=st["leaves"], hits=len(hits), seconds=round(secs, 4), device_seconds=round(dev_s, 4),
               groups_per_s=round(groups / secs, 2), tasks_per_s=round(st["tasks"] / secs, 1),
               ms_per_group=round(1e3 * secs / groups, 3),
               bundle_hours=round((1 << 20) * secs / groups / 3600, 2),
               roofline=dict(bound="valu", unit="Tops/s", peak=round(PEAK_VALU_OPS / 1e12, 2),
                             achieved=round(work / dev_s / 1e12, 3), frac=round(work / dev_s / PEAK_VALU_OPS, 4),
                             kernel="k_sv_lift_groups (with k_sv_regions and k_sv_pool in the same device time)",
                             work_unit="candidate = %d int32 lane-ops per lifted bit x 28 bits, + %d per full state"
                                       % (SOLVER_OPS_PER_LEVEL, SOLVER_OPS_PER_LEAF)),
               cpu_baseline=cpu)
    if cpu:
        out["gpu_over_cpu_tasks"] = round(out["tasks_per_s"] / cpu["value"], 1)
    return out


def full_c3_from_profiles():
    """The whole 2^32-structure-seed Deep Dive as one timed call (bench.py --dd-full), from its committed summary."""
    try:
        with open(os.path.join(REPO, "profiles", "deepdive_c3_full.json")) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def deep_dive_full(whf):
    """The Deep Dive workload in full: 2^32 structure seeds from the first SplitMix64(3) output, regions [-2,1]^2,
    room floor <= -55, every spot expanded to its 2^16 world seeds with the phase-1 chain and the swamp check (version
    1.20), as one whf_deepdive call, wall-timed."""
    s0 = splitmix_first(3) & ((1 << 48) - 1)
    piece = 1 << 26    # calls of 2^26 structure seeds, so that progress shows between them
    keys = ("starts", "spots", "expanded_seeds", "accepted", "biome_rejected", "overflow")
    st = dict.fromkeys(keys, 0)
    finds = []
    with whf.Context(0) as ctx:
        ctx.deepdive(s0, 1 << 20, DD_RECT, DD_FLOOR)   # warm-up
        t0 = time.perf_counter()
        for k in range(DD_TOTAL // piece):
            _, f, s = ctx.deepdive((s0 + k * piece) & ((1 << 48) - 1), piece, DD_RECT, DD_FLOOR, expand=True,
                                   max_height=-50, spot_capacity=1 << 21, find_capacity=1 << 16)
            finds += f
            for key in keys:
                st[key] += int(s[key])
            if k % 8 == 7:
                note("full Deep Dive: %d of %d calls, %.1f s" % (k + 1, DD_TOTAL // piece, time.perf_counter() - t0))
        secs = time.perf_counter() - t0
    note("full Deep Dive of 2^32 structure seeds: %.1f s" % secs)
    return dict(workload="2^32 structure seeds S = first SplitMix64(3) output + i (mod 2^48), hut regions [-2,1]^2, "
                         "room floor <= -55, spots expanded to 2^16 world seeds (phase-1 chain + swamp check, 1.20), "
                         "as 64 whf_deepdive calls of 2^26",
                structure_seeds=DD_TOTAL, seconds=round(secs, 2), structure_seeds_per_s=round(DD_TOTAL / secs, 1),
                starts=st["starts"], spots=st["spots"], world_seeds=st["expanded_seeds"],
                phase1_survivors=st["accepted"], biome_rejected=st["biome_rejected"], finds=len(finds),
                finds_list=[list(f) for f in finds[:64]], overflow=st["overflow"])


WORLD_RECT = (-58594, 58594, -58594, 58594)   # hunt.py:72 WORLD: hut regions of the +-30M world, [low, high)
# docs/HUNT54.txt:19-20 (and hunt.py:86 SCAN_REFERENCE): whole-world seeds with one confirmed hut each at the -54 hunt
WORLD_SEEDS = [(1473884015873122059, (-15792, -54, -17168)), (1291329195661654798, (2137600, -54, 26493184))]


def world_seeds(whf):
    """The reference's per-seed workload: one seed over the whole world (1.37e10 hut regions), phase 1 and lysh's hut
    verdict at -54 or-lower, for the two seeds its docs give with their confirmed huts; one scan per seed and precision
    after a warm-up. The reference quotes about 9.7 s per such seed for its optimised scanner on a 188-SM workstation
    GPU and 58.4 s on a 20-SM laptop GPU (docs/HUNT54.txt:10-16)."""
    import numpy as np
    vf = whf.version_flags("1.20")
    tree = whf.biome_tree("1.20")
    regions = (WORLD_RECT[1] - WORLD_RECT[0]) * (WORLD_RECT[3] - WORLD_RECT[2])
    secs = {}
    found = True
    with whf.Context(0) as ctx:
        ctx.scan((-8, 8, -8, 8), seeds=np.array([1], dtype=np.int64))   # warm-up
        for prec in (whf.PREC_FP64, whf.PREC_FP32_GATE):
            t = 0.0
            for seed, hut in WORLD_SEEDS:
                t0 = time.perf_counter()
                hits, st = ctx.scan(WORLD_RECT, seeds=np.array([seed], dtype=np.int64), max_height=-50, vflags=vf,
                                    capacity=1 << 12, precision=prec, phase2=(whf.P2_OR_LOWER, -54, tree))
                t += time.perf_counter() - t0
                found = found and [(h[3] * 16, h[5], h[4] * 16) for h in hits] == [hut]
            secs["fp64" if prec == whf.PREC_FP64 else "fp32_gate"] = t / len(WORLD_SEEDS)
    best = min(secs.values())
    note("whole-world seeds: %.3f s per seed" % best)
    return dict(workload="one seed over the whole world (hut regions [-58594, 58594)^2, 1.373e10 regions), phase 1 "
                         "at maxHeight -50, then lysh's hut verdict at -54 or-lower (1.19-26.1 rules); the two seeds "
                         "of docs/HUNT54.txt with their confirmed huts",
                seeds=[s for s, _ in WORLD_SEEDS], regions_per_seed=regions,
                seconds_per_seed={k: round(v, 3) for k, v in secs.items()},
                checks_per_s=round(regions / best, 1), known_huts_found=found,
                reference_seconds_per_seed={"188-SM workstation GPU (docs/HUNT54.txt:16)": 9.7,
                                            "20-SM laptop GPU (docs/HUNT54.txt:10-12)": 58.4})


def deep_dive(whf, sweep_n, expand_n, cpu_seconds):
    """Deep Dive: room-floor sweep over sweep_n structure seeds, and the 2^16 expansion of the spots of the first
    expand_n; both from the first SplitMix64(3) output, hut regions [-2,1]^2, room floor <= -55, version 1.20."""
    s0 = splitmix_first(3) & ((1 << 48) - 1)
    with whf.Context(0) as ctx:
        ctx.deepdive(s0, 1 << 20, DD_RECT, DD_FLOOR)   # warm-up
        t0 = time.perf_counter()
        spots, _, st = ctx.deepdive(s0, sweep_n, DD_RECT, DD_FLOOR, spot_capacity=1 << 21)
        t_sweep = time.perf_counter() - t0
        note("deep dive sweep of %d structure seeds: %.3f s" % (sweep_n, t_sweep))
        t0 = time.perf_counter()
        xspots, finds, xs = ctx.deepdive(s0, expand_n, DD_RECT, DD_FLOOR, expand=True, max_height=-50,
                                         spot_capacity=1 << 21)
        t_exp_total = time.perf_counter() - t0
        note("deep dive expansion of %d structure seeds: %.3f s" % (expand_n, t_exp_total))
    with whf.Context(0, profile=True) as ctx:
        _, _, ps = ctx.deepdive(s0, min(sweep_n, 1 << 26), DD_RECT, DD_FLOOR, spot_capacity=1 << 21)
        _, _, px = ctx.deepdive(s0, expand_n, DD_RECT, DD_FLOOR, expand=True, max_height=-50, spot_capacity=1 << 21)
    sweep_rate = sweep_n / t_sweep
    # expansion time: the expanding call's wall time less its own sweep of expand_n structure seeds at the measured
    # sweep rate (the expansion's kernels run on three streams, so summing their event times would overcount)
    x_sec = t_exp_total - expand_n / sweep_rate
    x_rate = xs["expanded_seeds"] / x_sec if x_sec > 0 else None
    spots_per_seed = st["spots"] / sweep_n
    total_s = DD_TOTAL / sweep_rate + (DD_TOTAL * spots_per_seed * 65536 / x_rate if x_rate else 0)
    built = px["samplers_built"]
    x_ops = sum(built) * OPS_PER_SAMPLER
    return dict(
        workload="Deep Dive: structure seeds S = first SplitMix64(3) output + i (mod 2^48), hut regions [-2,1]^2, "
                 "mineshaft room floor <= -55, each spot expanded to its 2^16 world seeds with the phase-1 check "
                 "(version 1.20, maxHeight -50)",
        sweep=dict(structure_seeds=sweep_n, seconds=round(t_sweep, 4), structure_seeds_per_s=round(sweep_rate, 1),
                   starts=st["starts"], walks=st["walks"], spots=st["spots"], lcg_draws=st["draws"],
                   stage_ms=[round(v, 3) for v in ps["stage_ms"][:2]], stage_ms_structure_seeds=min(sweep_n, 1 << 26)),
        expansion=dict(structure_seeds=expand_n, spots=xs["spots"], world_seeds=xs["expanded_seeds"],
                       seconds_total=round(t_exp_total, 4), stage_seconds=round(x_sec, 4),
                       world_seeds_per_s=round(x_rate, 1) if x_rate else None, finds=len(finds),
                       phase1_survivors=xs["accepted"], biome_checked=xs["biome_checked"],
                       biome_rejected=xs["biome_rejected"],
                       funnel=xs["funnel"], tiers=xs["tiers"], samplers_built=built,
                       roofline=dict(bound="valu", unit="Tops/s", peak=round(PEAK_VALU_OPS / 1e12, 2),
                                     achieved=round(x_ops / x_sec / 1e12, 3) if x_sec > 0 else None,
                                     frac=round(x_ops / x_sec / PEAK_VALU_OPS, 4) if x_sec > 0 else None,
                                     work_unit="Perlin sampler built (all sets) = 259 Xoroshiro128++ draws x 25 "
                                               "int32 lane-ops; evaluations not counted")),
        full_workload=dict(structure_seeds=DD_TOTAL, projected_seconds=round(total_s, 1),
                           structure_seeds_per_s=round(DD_TOTAL / total_s, 1) if total_s else None,
                           measured=full_c3_from_profiles()),
        cpu_baseline=None if cpu_seconds <= 0 else deep_dive_cpu(s0, cpu_seconds),
    )


def serving_stream(n, seed, versions=("1.20",), precision=0):
    """The serving workload: n requests, sizes round(2^U(0,20)) seeds from SplitMix64(request index), radius
    uniform over {512, 1024, 2048, 4096} blocks -> regions +-(r // 512 + 1), threshold uniform over -50..-54 with
    'or lower' / 'exactly' at 50/50 (-54 is always 'or lower'), version 1.20. Phase 1 runs at maxHeight -50, and
    lysh's hut verdict applies the request's threshold to the survivors."""
    import numpy as np
    rng = np.random.default_rng(seed)
    sizes = np.rint(2.0 ** rng.uniform(0, 20, n)).astype(np.int64)
    radius = rng.choice([512, 1024, 2048, 4096], n)
    thr = rng.integers(-54, -49, n)
    exactly = (rng.random(n) < 0.5) & (thr != -54)
    reqs = []
    for i in range(n):
        r = int(radius[i]) // 512 + 1
        reqs.append(dict(rect=(-r, r, -r, r), splitmix=(i, int(sizes[i])), threshold=int(thr[i]),
                         exactly=bool(exactly[i]), checks=int(sizes[i]) * (2 * r) ** 2,
                         version=versions[i % len(versions)], precision=precision))
    return reqs


def serve(whf, reqs, rate, slots, chunk_checks, background=None, slots_mode=False):
    """Submits reqs at Poisson arrivals of `rate` per second (rate None: all at once), after an optional background
    Deep Dive (s_first, count); returns per-request events, the wall time from the first submit to the last
    request's completion, and the Deep Dive's event. slots_mode: the host-driven slot scheduler instead of the
    device-resident loop."""
    import threading
    import numpy as np
    rng = np.random.default_rng(12345)
    gaps = rng.exponential(1.0 / rate, len(reqs)) if rate else np.zeros(len(reqs))
    events = {}
    dd_event = None
    with whf.Server(0, slots=slots, chunk_checks=chunk_checks, slots_mode=slots_mode) as srv:
        lock = threading.Lock()
        ids = []
        dd_id = None
        if background:
            dd_id = srv.submit_deepdive(background[0], background[1], DD_RECT, floor_max=DD_FLOOR, expand=True,
                                        max_height=-50)

        def feeder():
            t0 = time.perf_counter()
            t = 0.0
            for q, g in zip(reqs, gaps):
                t += g
                d = t0 + t - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                with lock:
                    ids.append(srv.submit(q["rect"], capacity=1024, splitmix=q["splitmix"], max_height=-50,
                                          vflags=whf.version_flags(q["version"]), precision=q["precision"],
                                          phase2=(whf.P2_EXACTLY if q["exactly"] else whf.P2_OR_LOWER,
                                                  q["threshold"], whf.biome_tree(q["version"]))))

        th = threading.Thread(target=feeder)
        t0 = time.perf_counter()
        th.start()
        while len(events) < len(reqs) or (dd_id is not None and dd_event is None):
            with lock:
                evs = srv.poll(timeout_ms=0)
            for e in evs:
                if e["id"] == dd_id:
                    dd_event = e
                else:
                    events[e["id"]] = e
                    if len(events) == len(reqs):
                        wall = time.perf_counter() - t0
            if not evs:
                time.sleep(0.0005)
        th.join()
    return [events[i] for i in ids], wall, dd_event


def scan_rate(whf, reqs):
    """whf_scan on the same requests' seeds, grouped by rect and version (one scan per group, seed lists): the rate a
    one-request-per-launch scan reaches on this mix, with every group's seeds in one request."""
    import numpy as np
    groups = {}
    for q in reqs:
        groups.setdefault((q["rect"], q["version"], q["precision"]), []).append(q)
    lists = {k: np.concatenate([splitmix_seeds(q["splitmix"][0], q["splitmix"][1]) for q in qs])
             for k, qs in groups.items()}
    best = None
    with whf.Context(0) as ctx:
        for _ in range(2):
            t = time.perf_counter()
            n = 0
            for (rect, version, precision), seeds in lists.items():
                _, st = ctx.scan(rect, seeds=seeds, max_height=-50, vflags=whf.version_flags(version),
                                 precision=precision, capacity=1 << 16)
                n += st["checks"]
            w = time.perf_counter() - t
            best = max(best or 0.0, n / w)
    return best


def serving(whf, n, load, slots, chunk_checks, versions=("1.20",), background=None, slots_mode=False, precision=0):
    import numpy as np
    reqs = serving_stream(n, 4, versions, precision)
    mean_checks = float(np.mean([q["checks"] for q in reqs]))
    # capacity: a prefix of the same stream submitted at once (saturation), the better of two runs
    sat = reqs[: max(50, n // 20)]
    capacity = 0.0
    for _ in range(2):
        ev, wall, _ = serve(whf, sat, None, slots, chunk_checks, slots_mode=slots_mode)
        capacity = max(capacity, sum(e["stats"]["checks"] for e in ev) / wall)
    scan = scan_rate(whf, sat)
    rate = load * capacity / mean_checks
    note("serving capacity %.3g checks/s from %d requests at once" % (capacity, len(sat)))
    ev, wall, dd = serve(whf, reqs, rate, slots, chunk_checks, background, slots_mode=slots_mode)
    lat = np.array([e["done_s"] - e["submit_s"] for e in ev]) * 1e3
    big = np.array([q["splitmix"][1] >= 1 << 19 for q in reqs])
    checks = sum(e["stats"]["checks"] for e in ev)
    assert all(e["status"] == whf.OK for e in ev)
    out = dict(
        workload="serving: %d Seed Check requests, Poisson arrivals at %.0f%% of measured capacity, sizes "
                 "round(2^U(0,20)) SplitMix64(request index) seeds, radius {512,1024,2048,4096} blocks, "
                 "threshold -50..-54 or-lower/exactly, versions %s, %s%s" % (
                     n, 100 * load, "/".join(versions), "FP32 gate" if precision else "FP64",
                     ", with a background Deep Dive of %d structure seeds" % background[1] if background else ""),
        scheduler=("host slots: %d contexts, chunks of %d checks" % (slots, chunk_checks)) if slots_mode else
                  ("device loop: %d lanes, batches of at most %d checks" % (slots, chunk_checks)),
        slots=slots, chunk_checks=chunk_checks, capacity_checks_per_s=round(capacity, 1),
        scan_same_requests_checks_per_s=round(scan, 1), capacity_vs_scan=round(capacity / scan, 3),
        mean_request_checks=round(mean_checks, 1), arrival_rate_per_s=round(rate, 2),
        offered_checks_per_s=round(rate * mean_checks, 1), throughput_checks_per_s=round(checks / wall, 1),
        wall_s=round(wall, 3), requests=n, finds=sum(len(e["hits"]) for e in ev),
        latency_ms=dict(p50=round(float(np.percentile(lat, 50)), 3), p99=round(float(np.percentile(lat, 99)), 3),
                        max=round(float(lat.max()), 3),
                        p99_large=round(float(np.percentile(lat[big], 99)), 3) if big.any() else None,
                        large_requests=int(big.sum()), large="requests of >= 2^19 seeds"),
    )
    if background:
        assert dd["status"] == whf.OK
        dd_s = dd["done_s"] - dd["start_s"]
        last_request = max(e["done_s"] for e in ev)
        out["deep_dive"] = dict(structure_seeds=background[1], seconds=round(dd_s, 3),
                                structure_seeds_per_s=round(background[1] / dd_s, 1), spots=len(dd["spots"]),
                                expanded_seeds=dd["dd_stats"]["expanded_seeds"], finds=len(dd["finds"]),
                                ran_under_whole_stream=bool(dd["done_s"] >= last_request))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--seeds", type=int, default=1 << 26)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="use the library's default launch shapes")
    ap.add_argument("--tune-seeds", type=int, default=1 << 24, help="seeds of the autotuner's timing workload")
    ap.add_argument("--product-seeds", type=int, default=1024,
                    help="seeds of the product-mode CPU timing (one lysh process per seed)")
    ap.add_argument("--dd-seeds", type=int, default=1 << 28, help="structure seeds of the Deep Dive sweep")
    ap.add_argument("--dd-expand", type=int, default=1 << 22, help="structure seeds whose spots are expanded")
    ap.add_argument("--no-deep-dive", action="store_true")
    ap.add_argument("--no-world", action="store_true", help="skip the whole-world single-seed scans")
    ap.add_argument("--solver-groups", type=int, default=256,
                    help="groups of the Deep Dive solver's timing (0: none)")
    ap.add_argument("--dd-full", action="store_true",
                    help="only time the whole 2^32-structure-seed Deep Dive (minutes) and print its JSON")
    ap.add_argument("--serve-requests", type=int, default=10000)
    ap.add_argument("--serve-load", type=float, default=0.8)
    ap.add_argument("--serve-slots", type=int, default=4, help="lanes of the device loop")
    ap.add_argument("--serve-host-slots", type=int, default=6,
                    help="slots of the host-slot scheduler (6: its best capacity of 3 to 6 slots on this workload)")
    ap.add_argument("--serve-chunk", type=int, default=1 << 24, help="checks per chunk of the slot scheduler")
    ap.add_argument("--serve-batch", type=int, default=1 << 25, help="checks per batch of the device loop at most")
    ap.add_argument("--serve-only", choices=("device", "slots"),
                    help="only serve the first 150 requests of the serving stream at once with this scheduler, once "
                         "(a short run to profile), and print the rate")
    ap.add_argument("--mixed-dd-seeds", type=int, default=1 << 26,
                    help="structure seeds of the background Deep Dive in the mixed serving run (0: no mixed run)")
    args = ap.parse_args()
    if args.serve_only:
        whf = __graft_entry__.load_whf()
        reqs = serving_stream(3000, 4)[:150]
        mode = args.serve_only
        ev, wall, _ = serve(whf, reqs, None, args.serve_slots if mode == "device" else args.serve_host_slots,
                            args.serve_batch if mode == "device" else args.serve_chunk, slots_mode=(mode == "slots"))
        print(json.dumps(dict(scheduler=mode, requests=len(reqs), checks=sum(e["stats"]["checks"] for e in ev),
                              seconds=round(wall, 4), checks_per_s=round(sum(e["stats"]["checks"] for e in ev) / wall))))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each GPU would scan its own seed "
              "range independently); run with --gpus 1", file=sys.stderr)
        return 2

    whf = __graft_entry__.load_whf()
    if args.dd_full:
        print(json.dumps(deep_dive_full(whf)), flush=True)
        return 0
    tuning = None
    if not args.no_tune:
        tuning = autotune(args.tune_seeds)
        note("autotuned in %.1f s: %s" % (tuning["seconds"], tuning["chosen"]))
    rect = (-8, 8, -8, 8)
    vflags = whf.version_flags("1.20")
    kw = dict(splitmix=(2, args.seeds), max_height=-50, vflags=vflags, capacity=1 << 20,
              phase2=(whf.P2_OR_LOWER, THRESHOLD, whf.biome_tree("1.20")))
    timing = {}
    results = {}
    with whf.Context(0) as ctx:
        for prec in (whf.PREC_FP64, whf.PREC_FP32_GATE):
            for _ in range(args.warmup):
                ctx.scan(rect, precision=prec, **kw)
            note("warmup done (precision %d)" % prec)
            # whf_scan returns after synchronizing the library's stream, so the wall clock between calls brackets
            # exactly the device work of the K steps
            t0 = time.perf_counter()
            for _ in range(args.steps):
                hits, st = ctx.scan(rect, precision=prec, **kw)
            timing[prec] = (time.perf_counter() - t0) / args.steps
            note("%d timed steps (precision %d): %.3f s per step, %d finds" % (args.steps, prec, timing[prec], len(hits)))
            results[prec] = (hits, st)
        err = ctx.noise_error([int(x) for x in range(1, 257)], 4096)
        # phase 2 on its own: the step's phase-1 survivors evaluated in one call (throughput), and one hut per call
        # (the latency a small request sees)
        kw1 = {k: v for k, v in kw.items() if k != "phase2"}
        surv, _ = ctx.scan(rect, precision=whf.PREC_FP64, **kw1)
        huts = [(h[0], h[3], h[4]) for h in surv]
        tree = whf.biome_tree("1.20")
        ctx.hut_eval(huts[:8], tree, THRESHOLD, vflags)
        t0 = time.perf_counter()
        ctx.hut_eval(huts, tree, THRESHOLD, vflags)
        p2_batch_s = time.perf_counter() - t0
        singles = huts[:20]
        t0 = time.perf_counter()
        for h in singles:
            ctx.hut_eval([h], tree, THRESHOLD, vflags)
        p2_single_s = (time.perf_counter() - t0) / max(1, len(singles))
        note("phase 2 alone: %d huts in %.1f ms, %.2f ms for one" % (len(huts), p2_batch_s * 1e3, p2_single_s * 1e3))
    h64, s64 = results[whf.PREC_FP64]
    h32, last = results[whf.PREC_FP32_GATE]
    identical = h64 == h32 and all(s64[k] == last[k] for k in ("funnel", "tiers", "accepted", "checks", "phase2"))
    # the headline is the faster of the two bit-identical paths; both are reported
    best = min(timing, key=timing.get)
    secs = timing[best]
    checks = last["checks"]

    # profiled step: per-kernel-family device time from HIP events recorded on the stream each kernel runs on
    fp32 = best == whf.PREC_FP32_GATE
    with whf.Context(0, profile=True) as ctx:
        _, prof = ctx.scan(rect, precision=best, **kw)
        shape = ctx.tuning(vflags)
    fam = prof["kernel_ms"]
    n_batches = -(-args.seeds // (1 << 17))
    built = prof["samplers_built"]
    work = family_work(prof, fp32)
    rooflines = {}
    for name, ms in fam.items():
        if name not in work or ms <= 0:
            continue
        rooflines[name] = dict(roofline_of(work[name], ms / 1e3), launch_ms=round(ms / n_batches, 4))
    # the same step with every stage on one stream (no overlap): each kernel's own duration, the figure a
    # kernel-level optimisation moves; the pipelined figures above include the device the launch shares
    with whf.Context(0, profile=True, one_stream=True) as ctx:
        _, prof1 = ctx.scan(rect, precision=best, **kw)
    rooflines_alone = {}
    for name, ms in prof1["kernel_ms"].items():
        if name not in work or ms <= 0:
            continue
        rooflines_alone[name] = dict(roofline_of(work[name], ms / 1e3), launch_ms=round(ms / n_batches, 4))
    # the noise kernel: the erosion gate of the faster precision
    # the template instance whf_scan launches for this rect (16 regions wide: 4 seeds per tile) and shape; the plain
    # FP64 kernel and the 512-wide workgroups take one item per lane
    ilp = shape["ero_ilp"] if fp32 and shape["ero_wg"] == 256 else 1
    kernel = "k_erosion<4, %s, %d, %d>" % ("true" if fp32 else "false", shape["ero_wg"], ilp)
    evals = erosion_evals(prof)
    r = rooflines["erosion_gate"]
    ra = rooflines_alone["erosion_gate"]
    # achieved in TFLOP/s of the peak named: for the FP32-gated kernel an FP64 FLOP counts as 2 FP32 FLOP (the ratio
    # of the two published peaks), so achieved / peak = frac
    peak = PEAK_FP32 if fp32 else PEAK_FP64
    roofline = dict(bound="valu", achieved=round(r["frac"] * peak / 1e12, 3), peak=peak / 1e12,
                    unit="TFLOP/s (FP32-equivalent)" if fp32 else "TFLOP/s", frac=r["frac"],
                    traffic=traffic_from_profiles(kernel), kernel=kernel,
                    launch_ms=r["launch_ms"], evals_per_launch=evals // n_batches,
                    flop_per_eval={"fp32": FLOP_BLEND_PER_EVAL if fp32 else 0,
                                   "fp64": FLOP_COORD_PER_EVAL + (0 if fp32 else FLOP_BLEND_PER_EVAL)},
                    fp32_tflops=r["fp32_tflops"], fp64_tflops=r["fp64_tflops"], frac_fp32=r["frac_fp32"],
                    frac_fp64=r["frac_fp64"], launch_ms_alone=ra["launch_ms"], frac_alone=ra["frac"],
                    frac_fp32_alone=ra["frac_fp32"], frac_fp64_alone=ra["frac_fp64"],
                    lane_ops_per_eval=OPS_PER_Y0_EVAL_LANE,
                    lane_op_frac=round(evals * OPS_PER_Y0_EVAL_LANE / (fam["erosion_gate"] / 1e3) / PEAK_VALU_OPS, 4),
                    work_unit="y=0 Perlin evaluation = %d FLOP corner blend (%s) + %d FLOP coordinates and "
                              "accumulation (FP64); evaluations per launch from the tier histogram; frac = FP32 "
                              "FLOP/s / 157.3 T + FP64 FLOP/s / 78.6 T" % (FLOP_BLEND_PER_EVAL,
                                                                           "FP32" if fp32 else "FP64",
                                                                           FLOP_COORD_PER_EVAL))
    # the largest kernel by device time, the sampler build (three sets), on its own line
    bms = sum(fam[k] for k in ("build_erosion", "build_climate", "build_ladder"))
    bms1 = sum(prof1["kernel_ms"][k] for k in ("build_erosion", "build_climate", "build_ladder"))
    bwork = (0, 0, sum(built) * OPS_PER_SAMPLER)
    rb = roofline_of(bwork, bms / 1e3)
    roofline_build = dict(bound="valu", kernel="k_build (sets E, C, L)", unit="Tops/s",
                          peak=round(PEAK_VALU_OPS / 1e12, 2), achieved=rb["int32_tops"], frac=rb["frac"],
                          frac_alone=roofline_of(bwork, bms1 / 1e3)["frac"], samplers=built,
                          launch_ms=round(bms / n_batches / 3, 4), ms_per_step=round(bms, 3),
                          ms_per_step_alone=round(bms1, 3),
                          traffic={k: traffic_from_profiles("k_build (set %s)" % k) for k in "ECL"},
                          work_unit="Perlin sampler built = 259 Xoroshiro128++ draws x 25 int32 lane-ops")
    # the whole step: all counted work over the step's wall time. A batch's gates 2 and 3 run beside the next
    # batch's builds and erosion gate, so a kernel's launch time includes the device it shares; this is the figure
    # the overlap improves.
    step_work = tuple(sum(w[i] for w in work.values()) for i in range(3))
    step_roofline = dict(roofline_of(step_work, secs), families=sorted(work), bound="valu")

    note("profiled step done")
    host = host_cpus()
    cpu = None if args.no_cpu else cpu_baseline(min(args.seeds, 1 << 20), args.cpu_seconds, host)
    note("cpu baseline done")
    parity = None
    if cpu:
        parity = cpu_parity(whf, cpu, (whf.PREC_FP64, whf.PREC_FP32_GATE))
        note("cpu parity: %s over %d seeds" % (parity["equal"], parity["seeds"]))
    product = None if args.no_cpu else cpu_product_mode(args.product_seeds, host["threads"])
    note("cpu product mode done")
    # serving: the device-resident loop (the default scheduler), then the host-driven slots on the same stream
    srv = srv_slots = None
    if args.serve_requests > 0:
        srv = serving(whf, args.serve_requests, args.serve_load, args.serve_slots, args.serve_batch)
        note("serving (device loop) done")
        srv_slots = serving(whf, args.serve_requests, args.serve_load, args.serve_host_slots, args.serve_chunk,
                            slots_mode=True)
        note("serving (slots) done")
    mixed = None
    if args.serve_requests > 0 and args.mixed_dd_seeds > 0:
        # three game versions with three different biome trees, FP32-gated noise, a background Deep Dive long
        # enough to run under the whole request stream
        mixed = serving(whf, args.serve_requests, args.serve_load, args.serve_slots, args.serve_batch,
                        versions=("1.18", "1.20", "1.21.5"), precision=whf.PREC_FP32_GATE,
                        background=(splitmix_first(3) & ((1 << 48) - 1), args.mixed_dd_seeds))
        note("mixed serving done")
    world = None if args.no_world else world_seeds(whf)
    dd = None if args.no_deep_dive else deep_dive(whf, args.dd_seeds, args.dd_expand,
                                                  0 if args.no_cpu else min(10.0, args.cpu_seconds))
    sol = None if args.solver_groups <= 0 else solver(whf, args.solver_groups,
                                                      0 if args.no_cpu else min(10.0, args.cpu_seconds), host)
    line = {
        "metric": METRIC, "value": round(checks / secs, 1), "unit": "checks/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(secs * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if best == whf.PREC_FP64 else "f32-gate+f64", "data": "synthetic",
        "config": {"workload": "Seed Check at scale: %d SplitMix64(2) seeds x regions [-8,7]^2 (256 per seed), "
                               "version 1.20, threshold %d or-lower (phase-1 maxHeight -50, then lysh's hut verdict "
                               "on the survivors); value = the faster of the plain FP64 chain and the FP32-gated "
                               "erosion stage (identical finds)" % (args.seeds, THRESHOLD),
                   "seeds": args.seeds, "regions_per_seed": 256, "version": "1.20", "threshold": THRESHOLD,
                   "mode": "or-lower", "precision": "fp64" if best == whf.PREC_FP64 else "fp32-gate"},
        "finds": len(h64), "phase1_survivors": last["accepted"],
        "phase2": {"evaluated": last["phase2"][0], "accepted": last["phase2"][1],
                   "rejected_biome": last["phase2"][2], "rejected_y": last["phase2"][3],
                   "rejected_flood": last["phase2"][4], "ms_per_step": round(last["phase2_ms"], 3),
                   "ms_per_hut": round(last["phase2_ms"] / max(1, last["phase2"][0]), 4),
                   "batch": {"huts": len(huts), "ms": round(p2_batch_s * 1e3, 3),
                             "ms_per_hut": round(p2_batch_s * 1e3 / max(1, len(huts)), 4)},
                   "single_hut_ms": round(p2_single_s * 1e3, 3),
                   "cpu_reference_ms_per_hut": cpu["phase2_ms_per_hut"] if cpu else None},
        "fp64_value": round(checks / timing[whf.PREC_FP64], 1), "fp64_ms_per_step": round(timing[whf.PREC_FP64] * 1e3, 3),
        "fp32_gate_value": round(checks / timing[whf.PREC_FP32_GATE], 1),
        "fp32_gate_ms_per_step": round(timing[whf.PREC_FP32_GATE] * 1e3, 3),
        "fp32_gate_over_fp64": round(timing[whf.PREC_FP64] / timing[whf.PREC_FP32_GATE], 3),
        "finds_identical_fp32_vs_fp64": identical,
        "fp64_rechecks": last["fp64_rechecks"],
        "fp32_error": {"stage": "erosion (8 octave samplers, y=0)", "max_abs_per_octave": err,
                       "bound_used": whf.fp32_bound()},
        "roofline": roofline, "roofline_build": roofline_build, "step_roofline": step_roofline,
        "cpu_baseline": cpu and {k: v for k, v in cpu.items() if k != "finds"},
        "cpu_parity": parity, "cpu_product_mode": product, "tuning": tuning,
        "funnel": last["funnel"], "tiers": last["tiers"], "accepted": last["accepted"],
        "kernel_ms": {k: round(v, 3) for k, v in fam.items()}, "rooflines": rooflines,
        "rooflines_alone": rooflines_alone, "samplers_built": built,
        "device_ms": round(last["device_ms"], 3),
        "latency_ms": srv["latency_ms"] if srv else None,
        "serving": srv,
        "serving_slots": srv_slots,
        "serving_mixed": mixed,
        "world_seed": world,
        "deep_dive": dd,
        "deep_dive_solver": sol,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
