"""Continuous batching (whf_server_* through the C ABI): every request's hits and counters in the mix equal those of
the same request scanned alone with whf_scan, whatever the batching, arrival order or overlap with other requests.
Each test runs against both schedulers: the device-resident loop (the default: batches packed on the device across
requests, per-request counters and survivor queues on the device) and the host-driven slots (WHF_SERVE_SLOTS).
"""
import time

import numpy as np
import pytest

import oracle_ctypes as O
from product import whf

pytestmark = pytest.mark.gpu

KEYS = ("checks", "accepted", "reached_ladder", "funnel", "tiers", "samplers_built", "fp64_rechecks")
MODES = ("device", "slots")


def server(mode, **kw):
    return whf.Server(0, slots_mode=(mode == "slots"), **kw)


def request_mix(n, seed, max_seeds_log2=14):
    """Requests of log-uniform size, mixed rects, versions, seed sources and precisions."""
    rng = np.random.default_rng(seed)
    versions = ["1.18", "1.20", "1.21", "26.2"]
    out = []
    for i in range(n):
        size = int(round(2 ** rng.uniform(0, max_seeds_log2)))
        r = int(rng.choice([2, 3, 5]))
        rect = (-r, r, -r, r)
        kw = dict(vflags=whf.version_flags(versions[i % 4]), max_height=-50,
                  precision=whf.PREC_FP32_GATE if i % 5 == 0 else whf.PREC_FP64)
        mode = i % 3
        if mode == 0:
            kw["splitmix"] = (1000 + i, size)
        elif mode == 1:
            kw["counter"] = (77 * i, size, 5)
        else:
            kw["seeds"] = O.seeds_splitmix(5000 + i, size)
        out.append((rect, kw))
    return out


def alone(reqs):
    res = []
    with whf.Context(0) as ctx:
        for rect, kw in reqs:
            hits, st = ctx.scan(rect, capacity=1 << 16, **kw)
            res.append((hits, st))
    return res


def check_same(ev, ref):
    hits, st = ref
    assert ev["status"] == whf.OK
    assert ev["hits"] == hits
    for k in KEYS:
        assert ev["stats"][k] == st[k], k
    assert ev["submit_s"] <= ev["start_s"] <= ev["done_s"]


@pytest.mark.timeout(600)
@pytest.mark.parametrize("mode", MODES)
def test_mix_equals_alone_all_at_once(mode):
    reqs = request_mix(48, 1)
    print("scanning %d requests alone" % len(reqs), flush=True)
    ref = alone(reqs)
    print("serving them together (%s)" % mode, flush=True)
    with server(mode, slots=4, chunk_checks=1 << 20) as srv:
        ids = [srv.submit(rect, capacity=1 << 16, **kw) for rect, kw in reqs]
        got = {}
        while len(got) < len(ids):
            for e in srv.poll(timeout_ms=60000):
                got[e["id"]] = e
    for rid, r in zip(ids, ref):
        check_same(got[rid], r)


@pytest.mark.timeout(600)
@pytest.mark.parametrize("mode", MODES)
def test_mix_with_phase2_equals_alone(mode):
    """Requests with lysh's hut verdict (or-lower and exactly, several thresholds and versions) at a phase-1 height
    that leaves survivors: finds with their floors and the phase-2 counters equal whf_scan alone."""
    rng = np.random.default_rng(9)
    versions = ["1.18", "1.20", "1.21", "26.2"]
    reqs = []
    for i in range(16):
        v = versions[i % 4]
        mode = whf.P2_EXACTLY if i % 2 else whf.P2_OR_LOWER
        thr = int(rng.choice([60, 64, 70, 75]))
        reqs.append(((-8, 8, -8, 8), dict(splitmix=(4242 + i, int(round(2 ** rng.uniform(9, 13)))), max_height=60,
                                          vflags=whf.version_flags(v), phase2=(mode, thr, whf.biome_tree(v)))))
    print("scanning %d phase-2 requests alone" % len(reqs), flush=True)
    ref = alone(reqs)
    assert sum(st["phase2"][0] for _, st in ref) > 20, "too few survivors to exercise phase 2"
    print("serving them together (%s)" % mode, flush=True)
    with server(mode, slots=4, chunk_checks=1 << 18) as srv:
        ids = [srv.submit(rect, capacity=1 << 16, **kw) for rect, kw in reqs]
        got = {}
        while len(got) < len(ids):
            for e in srv.poll(timeout_ms=60000):
                got[e["id"]] = e
    for rid, r in zip(ids, ref):
        check_same(got[rid], r)
        assert got[rid]["stats"]["phase2"] == r[1]["phase2"]


@pytest.mark.timeout(600)
@pytest.mark.parametrize("mode", MODES)
def test_mix_equals_alone_staggered_arrivals(mode):
    """Requests arrive while others run (one big request first, then a stream of small ones)."""
    reqs = [((-9, 9, -9, 9), dict(splitmix=(31337, 1 << 20), max_height=-50))] + request_mix(24, 2, 10)
    ref = alone(reqs)
    with server(mode, slots=3, chunk_checks=1 << 22) as srv:
        ids = []
        for rect, kw in reqs:
            ids.append(srv.submit(rect, capacity=1 << 16, **kw))
            time.sleep(0.002)
        got = {}
        while len(got) < len(ids):
            for e in srv.poll(timeout_ms=60000):
                got[e["id"]] = e
    for rid, r in zip(ids, ref):
        check_same(got[rid], r)
    # the small requests did not wait for the big one to finish
    big_done = got[ids[0]]["done_s"]
    assert sum(got[i]["done_s"] < big_done for i in ids[1:]) > 0


@pytest.mark.parametrize("mode", MODES)
def test_empty_request_and_sink_overflow(mode):
    with server(mode, slots=2) as srv:
        a = srv.submit((0, 1, 0, 1), splitmix=(1, 0))
        b = srv.submit((-79, 79, -79, 79), capacity=0, seeds=np.array([1473884015873122059], dtype=np.int64),
                       max_height=-50)
        got = {}
        while len(got) < 2:
            for e in srv.poll(timeout_ms=60000):
                got[e["id"]] = e
    assert got[a]["status"] == whf.OK and got[a]["stats"]["checks"] == 0
    assert got[b]["status"] == whf.EOVERFLOW and got[b]["stats"]["accepted"] >= 1
    assert got[b]["stats"]["checks"] == 158 * 158


@pytest.mark.timeout(600)
@pytest.mark.parametrize("mode", MODES)
def test_background_deepdive_and_requests_equal_alone(mode):
    """A background Deep Dive (4 chunks, the last partial) shares the slots with Seed Check requests of versions
    1.18, 1.20 and 1.21: every result equals its run alone."""
    s0 = O.splitmix(3, 0) & ((1 << 48) - 1)
    count = 3 * (1 << 20) + 12345
    rect = (-2, 2, -2, 2)
    reqs = []
    for i, v in enumerate(["1.18", "1.20", "1.21"] * 8):
        reqs.append(((-3, 3, -3, 3), dict(splitmix=(900 + i, 1 + 97 * i), max_height=-50,
                                          vflags=whf.version_flags(v))))
    ref = alone(reqs)
    with whf.Context(0) as ctx:
        dd_spots, dd_finds, dd_st = ctx.deepdive(s0, count, rect, floor_max=-55, expand=True, max_height=-50,
                                                 spot_capacity=1 << 16)
    with server(mode, slots=3) as srv:
        did = srv.submit_deepdive(s0, count, rect, floor_max=-55, expand=True, max_height=-50)
        ids = []
        for rect_i, kw in reqs:
            ids.append(srv.submit(rect_i, capacity=1 << 16, **kw))
            time.sleep(0.003)
        got = {}
        while len(got) < len(ids) + 1:
            for e in srv.poll(timeout_ms=60000):
                got[e["id"]] = e
    for rid, r in zip(ids, ref):
        check_same(got[rid], r)
    d = got[did]
    assert d["status"] == whf.OK
    assert d["spots"] == dd_spots
    assert d["finds"] == dd_finds
    for k in ("structure_seeds", "starts", "walks", "spots", "expanded_seeds", "accepted", "funnel", "tiers",
              "biome_checked", "biome_rejected"):
        assert d["dd_stats"][k] == dd_st[k], k
    # the interactive requests were not held behind the whole Deep Dive
    assert min(got[i]["done_s"] for i in ids) < d["done_s"]


@pytest.mark.timeout(300)
@pytest.mark.parametrize("mode", MODES)
def test_close_with_queued_deepdive_returns_without_spinning(mode):
    """Closing a server while a background Deep Dive still has chunks queued and every Deep Dive place is taken
    (3 slots, at most 1 Deep Dive chunk at a time): close finishes the Deep Dive and returns, and the two idle slots
    sleep meanwhile instead of spinning on the scheduler's lock (the process's CPU time over the close stays below
    twice its wall time; a thread waiting on the device may spin, two more spinning threads would pass it)."""
    s0 = O.splitmix(3, 0) & ((1 << 48) - 1)
    srv = server(mode, slots=3)
    srv.submit_deepdive(s0, 16 << 18, (-2, 2, -2, 2), floor_max=-55, expand=True, max_height=-50)
    time.sleep(0.05)
    t0, c0 = time.perf_counter(), time.process_time()
    srv.close()
    wall, cpu = time.perf_counter() - t0, time.process_time() - c0
    print("close: %.3f s wall, %.3f s cpu" % (wall, cpu), flush=True)
    assert wall > 0.05, "the Deep Dive finished before close; the case was not exercised"
    assert cpu < 2.0 * wall + 0.05


@pytest.mark.timeout(600)
def test_device_loop_survivor_queue_overflow_falls_back_to_scan():
    """A phase-2 request with more phase-1 survivors than the device loop reserves for it (4096) cannot be judged from
    its queue: the loop scans it again alone, so its finds and counters still equal whf_scan's. Small requests beside
    it are served from the queue."""
    big = ((-8, 8, -8, 8), dict(splitmix=(4343, 1 << 17), max_height=60, vflags=whf.version_flags("1.20") | 8,
                                phase2=(whf.P2_OR_LOWER, 80, whf.biome_tree("1.20"))))
    reqs = [big] + request_mix(6, 5, 8)
    ref = alone(reqs)
    assert ref[0][1]["accepted"] > 4096, ref[0][1]["accepted"]
    with server("device") as srv:
        ids = [srv.submit(rect, capacity=1 << 16, **kw) for rect, kw in reqs]
        got = {}
        while len(got) < len(ids):
            for e in srv.poll(timeout_ms=60000):
                got[e["id"]] = e
    for rid, r in zip(ids, ref):
        check_same(got[rid], r)
    assert got[ids[0]]["stats"]["phase2"] == ref[0][1]["phase2"]


@pytest.mark.timeout(300)
def test_device_loop_packs_requests_into_shared_batches():
    """64 one-seed requests of mixed rects, versions and precisions submitted at once complete from far fewer
    batches than requests: their work is packed together (each still equal to its scan alone)."""
    rng = np.random.default_rng(3)
    reqs = []
    for i in range(64):
        r = int(rng.choice([2, 3, 5, 9]))
        reqs.append(((-r, r, -r, r), dict(splitmix=(7000 + i, 1 + i % 3), max_height=-50,
                                           vflags=whf.version_flags(["1.18", "1.20", "26.2"][i % 3]),
                                           precision=whf.PREC_FP32_GATE if i % 2 else whf.PREC_FP64)))
    ref = alone(reqs)
    with server("device") as srv:
        ids = [srv.submit(rect, capacity=1 << 16, **kw) for rect, kw in reqs]
        got = {}
        while len(got) < len(ids):
            for e in srv.poll(timeout_ms=60000):
                got[e["id"]] = e
    for rid, r in zip(ids, ref):
        check_same(got[rid], r)
    # requests that completed at the same instant came out of one batch's record
    distinct_done = len({round(got[i]["done_s"], 6) for i in ids})
    print("64 requests completed at %d distinct times" % distinct_done, flush=True)
