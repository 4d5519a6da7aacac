"""The three-stage batch pipeline (a batch's gates 2 and 3 on two more streams beside the next batches' erosion gate,
three buffer sets) against the same work on one stream (WHF_CFG_ONE_STREAM): identical hits, floors and counters, for
Seed Check scans of several batches in both precisions and with phase 2, and for a Deep Dive expansion of many
batches of 16 spots. The one-stream path is the batch-after-batch order the other GPU tests check against the oracle.
"""
import pytest

import oracle_ctypes as O
from product import whf

pytestmark = pytest.mark.gpu

KEYS = ("checks", "accepted", "reached_ladder", "funnel", "tiers", "phase2")


@pytest.fixture(scope="module")
def ctxs():
    print("opening a pipelined and a one-stream context", flush=True)
    a, b = whf.Context(0), whf.Context(0, one_stream=True)
    yield a, b
    a.close()
    b.close()


@pytest.mark.parametrize("prec", [whf.PREC_FP64, whf.PREC_FP32_GATE])
def test_scan_batches(ctxs, prec):
    a, b = ctxs
    # 5 batches of 2^17 seeds and a partial one, 256 regions each
    kw = dict(splitmix=(77, 5 * (1 << 17) + 1234), max_height=-50, vflags=whf.version_flags("1.20"),
              capacity=1 << 18, precision=prec)
    print("scan: 6 batches, precision %d" % prec, flush=True)
    ha, sa = a.scan((-8, 8, -8, 8), **kw)
    hb, sb = b.scan((-8, 8, -8, 8), **kw)
    assert ha == hb and len(ha) > 0
    for k in KEYS:
        assert sa[k] == sb[k], k


def test_scan_batches_with_phase2(ctxs):
    a, b = ctxs
    kw = dict(splitmix=(78, 3 * (1 << 17)), max_height=60, vflags=whf.version_flags("1.20"), capacity=1 << 18,
              phase2=(whf.P2_OR_LOWER, 70, whf.biome_tree("1.20")))
    print("scan: 3 batches with phase 2", flush=True)
    ha, sa = a.scan((-8, 8, -8, 8), **kw)
    hb, sb = b.scan((-8, 8, -8, 8), **kw)
    assert ha == hb and sa["phase2"][0] > 100
    for k in KEYS:
        assert sa[k] == sb[k], k


def test_deepdive_expansion_batches(ctxs):
    a, b = ctxs
    s0 = O.splitmix(3, 0) & ((1 << 48) - 1)
    print("deep dive: 2^21 structure seeds with expansion", flush=True)
    ra = a.deepdive(s0, 1 << 21, (-2, 2, -2, 2), floor_max=-55, expand=True, max_height=-50, spot_capacity=1 << 16)
    rb = b.deepdive(s0, 1 << 21, (-2, 2, -2, 2), floor_max=-55, expand=True, max_height=-50, spot_capacity=1 << 16)
    assert ra[0] == rb[0] and len(ra[0]) > 32   # spots: more than two expansion batches
    assert ra[1] == rb[1]                        # finds
    for k in ("expanded_seeds", "accepted", "funnel", "tiers"):
        assert ra[2][k] == rb[2][k], k
