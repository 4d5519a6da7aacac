"""The benchmark's headline answer checked against the oracle: the Seed Check at 2^26 SplitMix64(2) seeds x regions
[-8, 7]^2, version 1.20, threshold -52 'or lower' (phase 1 at maxHeight -50, then lysh's hut verdict).

Every phase-1 survivor of the full workload gets the oracle's hut verdict (lysh_eval_hut, eval.c:121-226) on the
host cores; the GPU's verdict on each survivor (whf_hut_eval) must equal it field by field, and the finds and phase-2
counters of whf_scan with phase 2 (both precisions) must be exactly the huts the oracle accepts. The phase-1
survivor set itself is pinned by the reference scanner's funnel on a prefix of the same stream (test_gpu_parity.py)
and by the benchmark's cpu_parity.
"""
import os
from concurrent.futures import ThreadPoolExecutor

import pytest

import oracle_ctypes as O
from product import whf

pytestmark = pytest.mark.gpu

RECT = (-8, 8, -8, 8)
N_SEEDS = 1 << 26
THRESHOLD = -52
FIELDS = ("dir", "size_x", "size_z", "biome_ok", "winner", "occ_y", "wet", "flooded", "carve_skipped", "avg_y",
          "sum_h", "ok", "reject")
REJECT_INDEX = {O.WO_HUT_REJECT_BIOME: 2, O.WO_HUT_REJECT_Y: 3, O.WO_HUT_REJECT_FLOOD: 4}


def oracle_verdicts(huts, tree, max_y):
    """The oracle's hut verdicts, one world seed's terrain per task, on the host cores (the oracle's calls release
    the interpreter lock)."""
    def one(h):
        seed, cx, cz = h
        return O.Terrain(seed & 0xFFFFFFFFFFFFFFFF).hut_eval(tree, cx * 16, cz * 16, max_y)

    with ThreadPoolExecutor(min(16, os.cpu_count() or 1)) as ex:
        return list(ex.map(one, huts))


@pytest.mark.timeout(900)
def test_c2_headline_verdicts_match_oracle():
    tree = whf.biome_tree("1.20")
    vf = whf.version_flags("1.20")
    with whf.Context(0) as ctx:
        print("phase 1 over 2^26 seeds x 256 regions", flush=True)
        surv, st1 = ctx.scan(RECT, splitmix=(2, N_SEEDS), max_height=-50, vflags=vf, capacity=1 << 20)
        assert st1["accepted"] == len(surv) > 1000
        huts = [(h[0], h[3], h[4]) for h in surv]
        print("GPU verdicts of %d survivors" % len(huts), flush=True)
        gpu = ctx.hut_eval(huts, tree, THRESHOLD, vf)
        print("oracle verdicts of %d survivors" % len(huts), flush=True)
        ref = oracle_verdicts(huts, tree, THRESHOLD)
        bad = [(h, g, r) for h, g, r in zip(huts, gpu, ref) if any(g[f] != r[f] for f in FIELDS)]
        assert not bad, bad[:3]
        want = sorted((h[0], h[1], h[2], h[3], h[4], r["avg_y"]) for h, r in zip(surv, ref) if r["ok"])
        p2 = [len(huts), len(want), 0, 0, 0]
        for r in ref:
            if not r["ok"]:
                p2[REJECT_INDEX[r["reject"]]] += 1
        print("finds %s, phase-2 counters %s" % (want, p2), flush=True)
        assert want, "the workload has finds at floor <= -52"
        for prec in (whf.PREC_FP64, whf.PREC_FP32_GATE):
            print("whf_scan with phase 2, precision %d" % prec, flush=True)
            finds, st = ctx.scan(RECT, splitmix=(2, N_SEEDS), max_height=-50, vflags=vf, capacity=1 << 20,
                                 precision=prec, phase2=(whf.P2_OR_LOWER, THRESHOLD, tree))
            assert finds == want
            assert st["phase2"] == p2
            assert st["funnel"] == st1["funnel"] and st["tiers"] == st1["tiers"]
