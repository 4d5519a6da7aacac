"""whf_deepdive_noalloc, the Deep Dive sweep that allocates nothing in the call (its queues and counters live in a
static device array), against the existing entry point whf_deepdive with expand = 0 on the same requests: identical
spots (structure seed, region, chunk, carver seed, room floor, pieces) and counters, bit for bit, on two calls in a
row (the second smaller, so stale contents of the shared array would show), plus the sink-overflow and argument
errors. The sweep itself is pinned against the oracle by test_gpu_parity_deepdive.py; one small case here repeats
that check for this entry point.
"""
import pytest

import oracle_ctypes as O
from product import whf

pytestmark = pytest.mark.gpu

MASK48 = (1 << 48) - 1
KEYS = ("structure_seeds", "tests", "starts", "walks", "spots", "draws", "pieces")


def same_as_deepdive(ctx, s0, count, rect, floor_max):
    print("noalloc vs deepdive: S0 %d count %d rect %s floor %d" % (s0, count, rect, floor_max), flush=True)
    want_spots, _, want = ctx.deepdive(s0, count, rect, floor_max=floor_max, spot_capacity=1 << 18)
    got_spots, got = ctx.deepdive_noalloc(s0, count, rect, floor_max=floor_max, spot_capacity=1 << 18)
    assert not got["overflow"] and not want["overflow"]
    assert got_spots == want_spots
    for k in KEYS:
        assert got[k] == want[k], k
    return got_spots, got


def test_noalloc_sweep_equals_deepdive_twice():
    s0 = O.splitmix(3, 0) & MASK48
    with whf.Context(0) as ctx:
        # five sweep batches of the scratch's size (16 regions: 1,024,000 structure seeds per batch)
        spots1, st1 = same_as_deepdive(ctx, s0, 5 << 20, (-2, 2, -2, 2), -55)
        assert st1["starts"] > 0 and len(spots1) > 100
        # a smaller second call on another range, rect and floor line: nothing of the first call may remain
        spots2, st2 = same_as_deepdive(ctx, (s0 + (1 << 40)) & MASK48, 200_000, (-3, 1, -1, 2), -50)
        assert st2["starts"] > 0 and len(spots2) < len(spots1)


def test_noalloc_sweep_matches_oracle():
    s0 = O.splitmix(3, 0) & MASK48
    rect = (-2, 2, -2, 2)
    with whf.Context(0) as ctx:
        spots, st = ctx.deepdive_noalloc(s0, 8192, rect, floor_max=-55)
    ref = O.deepdive_sweep(s0, 8192, rect, floor_max=-55, cut=-12 + 55)
    assert st["starts"] == ref["starts"] and st["walks"] == ref["status"]
    assert spots == ref["spots"]


def test_noalloc_sink_overflow_and_errors():
    s0 = O.splitmix(3, 0) & MASK48
    rect = (-2, 2, -2, 2)
    with whf.Context(0) as ctx:
        full, st = ctx.deepdive_noalloc(s0, 1 << 20, rect, floor_max=-50, spot_capacity=1 << 18)
        assert len(full) > 8
        part, st2 = ctx.deepdive_noalloc(s0, 1 << 20, rect, floor_max=-50, spot_capacity=8)
        assert st2["overflow"] and st2["spots"] == len(full) and len(part) == 8
        assert set(part) <= set(full)
        for bad in (dict(expand=1), dict(rect=(0, 0, 0, 1)), dict(rect=(-40, 40, -40, 40))):
            req = whf.DdRequest(s0, 1024, *(bad.get("rect", rect)), -55, bad.get("expand", 0), -50, 0, 0, 0)
            sink = whf.DdSink((whf.DdSpot * 4)(), 4, 0)
            assert ctx.L.whf_deepdive_noalloc(ctx.h, req, sink, None) == whf.EINVAL, bad
