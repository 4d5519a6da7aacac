"""The Seed Check with lysh's phase 2 on the GPU (whf_scan with phase2, and phase1_scan --phase2-max-y), end to end.

Bar: the output set {(seed, chunk, floor)} equals the reference CPU scanner's `lysh scan --list` (the known answer of
the reference, tests/golden/lysh_funnels.json) and the oracle's phase 1 + hut evaluation on the same seeds; counts
are exact.
"""
import json
import os
import subprocess
import tempfile

import pytest

import oracle_ctypes as O
from product import whf

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(REPO, "tests", "golden")
KNOWN = 1473884015873122059


@pytest.fixture(scope="module")
def ctx():
    print("opening whf context on device 0", flush=True)
    c = whf.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def known():
    with open(os.path.join(GOLD, "lysh_funnels.json")) as f:
        return json.load(f)["known_answer"]


def test_known_answer_end_to_end(ctx, known):
    """lysh scan --list over +-79 regions (hunt.py:86): one hut, floor -54."""
    tree = whf.biome_tree(known["version"])
    vf = whf.version_flags(known["version"])
    print("scan +-79 regions with phase 2", flush=True)
    hits, st = ctx.scan(tuple(known["rect"]), seeds=[KNOWN], max_height=-50, vflags=vf,
                        phase2=(whf.P2_OR_LOWER, -50, tree))
    # lysh lists "hit rx rz x y z flooded" with rx, rz = the hut block / 32 (eval.c:125-127) and y = the floor
    want = [(KNOWN, x >> 4, z >> 4, y) for rx, rz, x, y, z, flooded in known["hits"]]
    assert [(h[0], h[3], h[4], h[5]) for h in hits] == want
    assert [(h[3] * 16 // 32, h[4] * 16 // 32) for h in hits] == [(rx, rz) for rx, rz, *_ in known["hits"]]
    assert st["phase2"][0] == st["accepted"] and st["phase2"][1] == len(want)
    assert sum(st["phase2"][1:]) == st["phase2"][0]
    exact, _ = ctx.scan(tuple(known["rect"]), seeds=[KNOWN], max_height=-50, vflags=vf,
                        phase2=(whf.P2_EXACTLY, -54, tree))
    assert exact == hits
    none, _ = ctx.scan(tuple(known["rect"]), seeds=[KNOWN], max_height=-50, vflags=vf,
                       phase2=(whf.P2_EXACTLY, -53, tree))
    assert none == []


def test_phase2_matches_oracle(ctx):
    """Phase-1 survivors of 8,192 seeds at maxHeight 60, then lysh's verdict at floor <= 70 (tree 3)."""
    seeds = O.seeds_splitmix(1, 8192)
    rect = (-16, 16, -16, 16)
    print("scan with phase 2", flush=True)
    hits, st = ctx.scan(rect, seeds=seeds, max_height=60, vflags=0, phase2=(whf.P2_OR_LOWER, 70, 3))
    ref = O.scan(seeds, rect, 60, 0)
    want = []
    rej = [0, 0, 0]
    for s, rx, rz, cx, cz in ref["hits"]:
        r = O.Terrain(s & 0xFFFFFFFFFFFFFFFF).hut_eval(3, cx * 16, cz * 16, 70)
        if r["ok"]:
            want.append((s, rx, rz, cx, cz, r["avg_y"]))
        else:
            rej[{O.WO_HUT_REJECT_BIOME: 0, O.WO_HUT_REJECT_Y: 1, O.WO_HUT_REJECT_FLOOD: 2}[r["reject"]]] += 1
    assert len(ref["hits"]) >= 10
    assert hits == sorted(want)
    assert st["phase2"] == [len(ref["hits"]), len(want)] + rej


def test_phase1_scan_cli_phase2(known):
    exe = os.path.join(REPO, "low-witch-finder_amd", "bin", "phase1_scan")
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "c.csv")
        r = subprocess.run([exe, "--seed", str(KNOWN), "--min-region-x", "-79", "--max-region-x", "79",
                            "--min-region-z", "-79", "--max-region-z", "79", "--max-height", "-50",
                            "--phase2-max-y", "-50", "--version", known["version"], "--out", out],
                           capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        with open(out) as f:
            rows = f.read().splitlines()
    assert rows[0] == "seed,chunkX,chunkZ,blockX,blockZ,gpuEst"
    assert rows[1:] == ["%d,%d,%d,%d,%d,%.3f" % (KNOWN, x >> 4, z >> 4, x, z, y) for _, _, x, y, z, _ in known["hits"]]
