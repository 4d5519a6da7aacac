"""Autotuner: measures the library's launch shapes (whf_tuning) on this device and writes tuned.json.

For each game version's rules (1.18, 1.20 and 1.21 by default; 1.20 and 1.21 share theirs), a greedy search times
whf_scan on a Seed Check workload of the benchmark's shape (2^24 SplitMix64(2) seeds x regions [-8, 7]^2, threshold
-52 or-lower) under each candidate: first the erosion gate's shape (items per lane, workgroup size, persistent grid),
then the grid of gates 2 and 3 with the best erosion shape. A candidate replaces the best so far only when it is faster
by more than the margin (default 2%): one shape's time varies by about that much from run to run, and a shape that
wins a small workload by less has lost the full one. Every candidate must give the default's exact answer
(finds, funnel, tier histogram), or the tuner stops. Then the server's chunk size: the smallest chunk whose serving
capacity (a fixed set of requests submitted at once) is within 5% of the best measured, since larger chunks raise
latency for little capacity. whf_open applies the device's entry when its "build" matches the library
(whf_build_id). hunt.py keeps a tuned configuration per GPU the same way (hunt.py:927-945, 959-1021).

    python tune.py [--out FILE] [--seeds N] [--versions 1.18,1.20,1.21] [--reps 2] [--precision fp32|fp64]
                   [--margin 0.02] [--no-serving] [--serving-requests N]

The default output is tuned.json beside libwhf_hip.so, where whf_open looks unless WHF_TUNED names another file;
entries of other devices already in the file are kept.
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import numpy as np  # noqa: E402

import whf  # noqa: E402

RECT = (-8, 8, -8, 8)
THRESHOLD = -52
# (ero_wg, ero_ilp, ero_wg_per_cu): the default first
ERO_CANDIDATES = [(256, 2, -1), (256, 1, -1), (512, 1, -1), (256, 2, 8), (256, 1, 10)]
GATE_CANDIDATES = [16, 8, 4]
CHUNK_CANDIDATES = [1 << 22, 1 << 23, 1 << 24, 1 << 25, 1 << 26]


def note(msg):
    print("[tune %.1fs] %s" % (time.perf_counter() - T0, msg), file=sys.stderr, flush=True)


T0 = time.perf_counter()


def time_scan(version, n_seeds, reps, precision, **tuning):
    """Best wall time of whf_scan under the given shapes, and its answer."""
    vf = whf.version_flags(version)
    tree = whf.biome_tree(version)
    kw = dict(splitmix=(2, n_seeds), max_height=-50, vflags=vf, capacity=1 << 20, precision=precision,
              phase2=(whf.P2_OR_LOWER, THRESHOLD, tree))
    with whf.Context(0) as ctx:
        ctx.set_tuning(vflags=vf, **tuning)
        ctx.scan(RECT, **dict(kw, splitmix=(2, 1 << 17)))   # warm-up
        best, answer = None, None
        for _ in range(reps):
            t0 = time.perf_counter()
            hits, st = ctx.scan(RECT, **kw)
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
            answer = (hits, st["funnel"], st["tiers"], st["phase2"])
    return best, answer


def tune_version(version, n_seeds, reps, precision, margin=0.02):
    base = whf.Context(0)
    vf = whf.version_flags(version)
    cur = base.tuning(vf)
    base.close()
    log = []
    ref_t, ref = time_scan(version, n_seeds, reps, precision, **cur)
    log.append(dict(shape=dict(cur), ms=round(ref_t * 1e3, 3)))
    note("%s: default %.1f ms" % (version, ref_t * 1e3))
    best_t, best = ref_t, dict(cur)
    for wg, ilp, per_cu in ERO_CANDIDATES:
        cand = dict(cur, ero_wg=wg, ero_ilp=ilp, ero_wg_per_cu=per_cu)
        if cand == cur:
            continue
        t, ans = time_scan(version, n_seeds, reps, precision, **cand)
        if ans != ref:
            raise SystemExit("tuning candidate %r changed the answer for %s" % (cand, version))
        log.append(dict(shape=dict(cand), ms=round(t * 1e3, 3)))
        note("%s: %r %.1f ms" % (version, (wg, ilp, per_cu), t * 1e3))
        if t < best_t * (1.0 - margin):
            best_t, best = t, cand
    for g in GATE_CANDIDATES:
        cand = dict(best, gate_wg_per_cu=g)
        if cand == best:
            continue
        t, ans = time_scan(version, n_seeds, reps, precision, **cand)
        if ans != ref:
            raise SystemExit("tuning candidate %r changed the answer for %s" % (cand, version))
        log.append(dict(shape=dict(cand), ms=round(t * 1e3, 3)))
        note("%s: gate %d/CU %.1f ms" % (version, g, t * 1e3))
        if t < best_t * (1.0 - margin):
            best_t, best = t, cand
    return best, dict(default_ms=round(ref_t * 1e3, 3), tuned_ms=round(best_t * 1e3, 3), candidates=log)


def serving_requests(n, seed):
    """n Seed Check requests of the serving workload's shape: log-uniform sizes up to 2^20 seeds, radius {512, 1024,
    2048, 4096} blocks, thresholds -50..-54."""
    rng = np.random.default_rng(seed)
    sizes = np.rint(2.0 ** rng.uniform(0, 20, n)).astype(np.int64)
    radius = rng.choice([512, 1024, 2048, 4096], n)
    thr = rng.integers(-54, -49, n)
    out = []
    for i in range(n):
        r = int(radius[i]) // 512 + 1
        out.append(((-r, r, -r, r), dict(splitmix=(i, int(sizes[i])), max_height=-50,
                                         phase2=(whf.P2_OR_LOWER, int(thr[i]), whf.biome_tree("1.20")))))
    return out


def serving_capacity(chunk, reqs):
    checks = sum(kw["splitmix"][1] * (r[1] - r[0]) * (r[3] - r[2]) for r, kw in reqs)
    with whf.Server(0, slots=4, chunk_checks=chunk) as srv:
        t0 = time.perf_counter()
        for rect, kw in reqs:
            srv.submit(rect, capacity=1024, **kw)
        got = 0
        while got < len(reqs):
            got += len(srv.poll(timeout_ms=60000))
        return checks / (time.perf_counter() - t0)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--out", default=os.path.join(HERE, "tuned.json"))
    ap.add_argument("--seeds", type=int, default=1 << 24)
    ap.add_argument("--versions", default="1.18,1.20,1.21")
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--precision", choices=("fp32", "fp64"), default="fp32")
    ap.add_argument("--margin", type=float, default=0.02, help="the least relative gain that replaces the best shape")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--serving-requests", type=int, default=200)
    a = ap.parse_args(argv)
    precision = whf.PREC_FP32_GATE if a.precision == "fp32" else whf.PREC_FP64
    with whf.Context(0) as ctx:
        device = ctx.device_name.split(" (")[0]
    versions = [v.strip() for v in a.versions.split(",") if v.strip()]
    layouts = {}
    for v in versions:   # one search per distinct rule layout; versions that share rules share the result
        layouts.setdefault(whf.version_flags(v), []).append(v)
    results = {}
    for vf, names in layouts.items():
        best, meas = tune_version(names[0], a.seeds, a.reps, precision, a.margin)
        results[vf] = (names, best, meas)
    # the entry's base shapes: those of the default rules (vflags 0) if tuned, else the first version's
    base_vf = 0 if 0 in results else next(iter(results))
    base = dict(results[base_vf][1])
    entry = {"build": whf.build_id()}
    entry.update({k: base[k] for k in whf.TUNING_FIELDS})
    versions_out = {}
    for vf, (names, best, meas) in results.items():
        diff = {k: best[k] for k in whf.TUNING_FIELDS if best[k] != base[k]}
        if diff:
            for nm in names:
                versions_out[nm] = diff
    if versions_out:
        entry["versions"] = versions_out
    measured = {"workload": "%d SplitMix64(2) seeds x regions [-8,7]^2, threshold %d or-lower, %s" % (
        a.seeds, THRESHOLD, a.precision), "per_version": {}}
    for vf, (names, best, meas) in results.items():
        measured["per_version"]["/".join(names)] = dict(meas, chosen={k: best[k] for k in whf.TUNING_FIELDS})
    if not a.no_serving:
        reqs = serving_requests(a.serving_requests, 7)
        caps = {}
        for ch in CHUNK_CANDIDATES:
            caps[ch] = serving_capacity(ch, reqs)
            note("serving chunk 2^%d: %.3g checks/s" % (ch.bit_length() - 1, caps[ch]))
        top = max(caps.values())
        entry["chunk_checks"] = min(ch for ch, c in caps.items() if c >= 0.95 * top)
        measured["serving_capacity_checks_per_s"] = {str(ch): round(c, 1) for ch, c in caps.items()}
    entry["measured"] = measured
    doc = {}
    if os.path.exists(a.out):
        try:
            with open(a.out) as f:
                doc = json.load(f)
        except (OSError, ValueError):
            doc = {}
    doc[device] = entry
    tmp = a.out + ".tmp"
    with open(tmp, "w") as f:
        json.dump(doc, f, indent=1)
    os.replace(tmp, a.out)
    note("wrote %s for %s" % (a.out, device))
    return device, entry


if __name__ == "__main__":
    dev, ent = main()
    print(json.dumps({dev: ent}))
