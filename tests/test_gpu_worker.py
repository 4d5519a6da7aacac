"""The `phase1_scan --worker` protocol as hunt.py's GpuScanWorker drives it (hunt.py:1276-1332), against the reference
engine's worker loop (engine/gpu/low54/phase1_scan.cu:1346-1387).

The worker starts in the run's output folder on line-buffered text pipes, prints ===WORKER_READY=== once its device
buffers exist, reads one request per stdin line (a UTF-8 BOM is stripped), answers ===SCAN_DONE rc=N=== after each
request and ===SCAN_ERROR: bad request line=== for a line it cannot parse, after which it keeps serving, and exits 0
at end of input. The candidate CSV's first three columns (the only ones hunt.py reads, hunt.py:1313-1317) must be the
oracle's phase-1 survivors.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle_ctypes as O
from product import whf

pytestmark = pytest.mark.gpu

KNOWN = 1473884015873122059


class Worker:
    def __init__(self, cwd):
        # stderr to a log file, as hunt.py does (gpu.log)
        self.log = open(os.path.join(cwd, "gpu.log"), "w")
        self.p = subprocess.Popen([whf.PHASE1_SCAN, "--worker"], stdin=subprocess.PIPE, stdout=subprocess.PIPE,
                                  stderr=self.log, text=True, bufsize=1, cwd=cwd)
        self.seen = []

    def wait(self, token):
        """The first stdout line starting with token; an ===SCAN_ERROR line or end of output fails the test."""
        for line in self.p.stdout:
            line = line.strip()
            self.seen.append(line)
            if line.startswith("===SCAN_ERROR"):
                return line
            if line.startswith(token):
                return line
        self.log.flush()
        with open(self.log.name) as f:
            raise AssertionError("worker output ended before %r: %r / %s" % (token, self.seen[-5:], f.read()))

    def send(self, line):
        self.p.stdin.write(line + "\n")
        self.p.stdin.flush()


def request(seed, rect):
    """The line GpuScanWorker.scan writes (hunt.py:1303-1304)."""
    return ("--seed %d --min-region-x %d --max-region-x %d --min-region-z %d --max-region-z %d --max-height -50 "
            "--out cands.csv" % ((seed,) + tuple(rect)))


def read_cands(path):
    """hunt.py's reader: skip the header, columns 0-2 of each row with at least 3 fields."""
    rows = []
    with open(path, encoding="utf-8") as fh:
        assert next(fh).strip() == "seed,chunkX,chunkZ,blockX,blockZ,gpuEst"
        for line in fh:
            p = line.strip().split(",")
            if len(p) >= 3:
                rows.append((int(p[0]), int(p[1]), int(p[2])))
    return sorted(rows)


def oracle_cands(seed, rect):
    ref = O.scan(np.array([seed], np.int64), rect, -50, whf.version_flags("1.19.x~26.1"), threads=8)
    return sorted((s, cx, cz) for s, rx, rz, cx, cz in ref["hits"])


@pytest.mark.timeout(300)
def test_worker_protocol_as_hunt_drives_it(tmp_path):
    w = Worker(str(tmp_path))
    try:
        print("waiting for ===WORKER_READY===", flush=True)
        assert w.wait("===WORKER_READY===") == "===WORKER_READY==="
        out = tmp_path / "cands.csv"
        # the reference's known-answer seed over +-79 regions (hunt.py:86)
        rect = (-79, 79, -79, 79)
        print("request: known-answer seed", flush=True)
        w.send(request(KNOWN, rect))
        assert w.wait("===SCAN_DONE") == "===SCAN_DONE rc=0==="
        want = oracle_cands(KNOWN, rect)
        assert (KNOWN, -15792 >> 4, -17168 >> 4) in want
        assert read_cands(out) == want
        os.remove(out)
        # the same request behind a UTF-8 byte-order mark
        print("request: BOM-prefixed line", flush=True)
        w.send("\ufeff" + request(KNOWN, rect))
        assert w.wait("===SCAN_DONE") == "===SCAN_DONE rc=0==="
        assert read_cands(out) == want
        os.remove(out)
        # a negative world seed (hunt.py writes Python ints with %d) with a phase-1 survivor at region (4, 11)
        neg = -3339079979306198497
        rect2 = (-16, 16, -16, 16)
        print("request: negative seed", flush=True)
        w.send(request(neg, rect2))
        assert w.wait("===SCAN_DONE") == "===SCAN_DONE rc=0==="
        got = read_cands(out)
        assert got == oracle_cands(neg, rect2)
        assert (neg, 128, 358) in got
        # malformed lines: an error line each, and the worker keeps serving
        for bad in ("--seed", "--seed 5 --min-region-x 3 --max-region-x 3 --min-region-z 0 --max-region-z 1",
                    "--bogus-flag 1"):
            print("request: malformed %r" % bad, flush=True)
            w.send(bad)
            assert w.wait("===SCAN_") == "===SCAN_ERROR: bad request line==="
        print("request after the errors", flush=True)
        w.send(request(KNOWN, rect))
        assert w.wait("===SCAN_DONE") == "===SCAN_DONE rc=0==="
        assert read_cands(out) == want
        # end of input: a clean exit
        w.p.stdin.close()
        assert w.p.wait(60) == 0
    finally:
        if w.p.poll() is None:
            w.p.kill()
            w.p.wait()
        w.log.close()


@pytest.mark.timeout(300)
def test_worker_async_requests_with_ids(tmp_path):
    """Request lines with --id K go to the library's serving loop while the worker reads on: three are in flight at
    once, each answered by ===SCAN_DONE id=K rc=0=== after its own CSV is written, and that CSV holds the oracle's
    survivors. A sequential line among them keeps its plain ===SCAN_DONE rc=0=== answer. At end of input the worker
    answers every pending request before it exits 0."""
    w = Worker(str(tmp_path))
    try:
        assert w.wait("===WORKER_READY===") == "===WORKER_READY==="
        neg = -3339079979306198497
        jobs = {"a": (KNOWN, (-79, 79, -79, 79)), "b": (neg, (-16, 16, -16, 16)), "c7": (KNOWN, (-20, 5, -3, 30))}
        for k, (seed, rect) in jobs.items():
            line = request(seed, rect).replace("--out cands.csv", "--out cands_%s.csv" % k)
            print("async request %s" % k, flush=True)
            w.send("--id %s %s" % (k, line))
        print("sequential request", flush=True)
        w.send(request(KNOWN, (-8, 8, -8, 8)))
        done, plain = {}, 0
        while len(done) < len(jobs) or plain < 1:
            line = w.wait("===SCAN_")
            assert not line.startswith("===SCAN_ERROR"), line
            if line.startswith("===SCAN_DONE id="):
                k, rc = line[len("===SCAN_DONE id="):-3].split(" rc=")
                done[k] = int(rc)
            else:
                assert line == "===SCAN_DONE rc=0===", line
                plain += 1
        assert done == {k: 0 for k in jobs}
        for k, (seed, rect) in jobs.items():
            assert read_cands(tmp_path / ("cands_%s.csv" % k)) == oracle_cands(seed, rect), k
        assert read_cands(tmp_path / "cands.csv") == oracle_cands(KNOWN, (-8, 8, -8, 8))
        # a request sent last, then end of input: answered before the exit
        w.send("--id last " + request(neg, (-16, 16, -16, 16)).replace("cands.csv", "cands_last.csv"))
        w.p.stdin.close()
        rest = [ln.strip() for ln in w.p.stdout]
        assert "===SCAN_DONE id=last rc=0===" in rest, rest[-5:]
        assert w.p.wait(60) == 0
        assert read_cands(tmp_path / "cands_last.csv") == oracle_cands(neg, (-16, 16, -16, 16))
    finally:
        if w.p.poll() is None:
            w.p.kill()
            w.p.wait()
        w.log.close()
