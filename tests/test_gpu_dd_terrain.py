"""The terrain stage of the reference's --spots mode (hut_lookup_v4_source.cu:1262-1292) on the GPU:
whf_dd_terrain through the C ABI and bin/deepdive_scan --spots with the command line hunt.py builds
(hunt.py:954-957).

Bar: integer work, so everything is exact. The filled spots equal the oracle's carver seed, start test and layout
walk; the finds equal those of the Deep Dive sweep with expansion over the same starts (whose expansion is pinned
against the oracle's phase-1 scan in test_gpu_parity_deepdive.py); the CLI's survivor file holds the reference's
header and one "seed,chunkX,chunkZ,layoutX,0,0" row per find.
"""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle_ctypes as O
from product import whf

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
M48 = (1 << 48) - 1
INT32_MIN = -(1 << 31)


@pytest.fixture(scope="module")
def ctx():
    print("opening whf context on device 0", flush=True)
    c = whf.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def mst():
    with open(os.path.join(GOLD, "mstree_vectors.json")) as f:
        return json.load(f)


def non_starts(n, seed):
    """(S, cx, cz) triples whose chunk is not a mineshaft start."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        s48 = int(rng.integers(0, 1 << 48, dtype=np.uint64))
        cx, cz = (int(v) for v in rng.integers(-40000, 40000, size=2))
        if not O.is_start(O.carver_seed(s48, cx, cz)):
            out.append((s48, cx, cz))
    return out


def test_spots_match_sweep_and_oracle(ctx):
    """The starts of a 4,096-seed sweep over 4 x 4 regions (every completed walk kept) plus 50 chunks that are not
    starts, shuffled: every spot filled as the oracle says, finds equal to the sweep's expansion."""
    s0 = O.splitmix(5, 0) & M48
    print("sweep", flush=True)
    spots, finds, st = ctx.deepdive(s0, 1 << 12, (-2, 2, -2, 2), floor_max=63, expand=True, max_height=-50,
                                    biome=False)
    assert len(spots) > 100 and st["walks"][2] == 0
    bad = non_starts(50, 7)
    order = [(s[0], s[3], s[4]) for s in spots] + bad
    np.random.default_rng(3).shuffle(order)
    print("terrain stage over %d spots" % len(order), flush=True)
    out, tfinds, tst = ctx.dd_terrain(order, max_height=-50)
    by_key = {(s[0], s[3], s[4]): s for s in spots}
    for (s48, cx, cz), got in zip(order, out):
        X = O.carver_seed(s48, cx, cz)
        assert got[:6] == (s48, cx >> 5, cz >> 5, cx, cz, X)
        if (s48, cx, cz) in by_key:
            assert got == by_key[(s48, cx, cz)]
            status, floor, pieces, _ = O.room_floor(X)
            assert (status, floor, pieces) == (0, got[6], got[7])
        else:
            assert got[6:] == (INT32_MIN, 0)
    assert tfinds == finds
    assert tst["structure_seeds"] == tst["tests"] == len(order)
    assert tst["starts"] == tst["spots"] == len(spots)
    assert tst["expanded_seeds"] == len(spots) << 16
    assert tst["funnel"] == st["funnel"] and tst["tiers"] == st["tiers"] and tst["accepted"] == st["accepted"]


def test_known_deep_huts_through_spots(ctx, mst):
    """The confirmed deep huts of the reference's data/finds.tsv, given as spots: each is a start with the
    recorded room floor and piece count, and its world seed is among the finds."""
    spots = [(f["seed"] & M48, f["cx"], f["cz"]) for f in mst["finds"] if f["start"]]
    out, finds, st = ctx.dd_terrain(spots, max_height=-50)
    want = {(f["seed"] & M48, f["cx"], f["cz"]): f for f in mst["finds"] if f["start"]}
    for got in out:
        f = want[(got[0], got[3], got[4])]
        assert (got[6], got[7]) == (f["room_floor"], f["pieces"])
    found = {(h[0], h[3], h[4]) for h in finds}
    for f in want.values():
        assert (f["seed"], f["cx"], f["cz"]) in found
    assert all(h[5] == want[(h[0] & M48, h[3], h[4])]["room_floor"] for h in finds)


def test_dd_terrain_rejects_bad_requests(ctx):
    with pytest.raises(whf.WhfError):
        ctx.dd_terrain([(1, 0, 0)], max_height=-500)
    with pytest.raises(whf.WhfError):
        ctx.dd_terrain([(1, 0, 0)], vflags=99)
    out, finds, st = ctx.dd_terrain([])
    assert out == [] and finds == [] and st["starts"] == 0


def run_cli(tmp_path, lines, extra=()):
    sp, out = tmp_path / "spots.txt", tmp_path / "surv.csv"
    sp.write_text("".join(l + "\n" for l in lines))
    # hunt.py's terrain command line (hunt.py:954-957, engine_args :1024-1029) with deepdive_scan in hut_lookup_v4's
    # place; the catalogue file is not read
    cmd = [whf.DEEPDIVE_SCAN, "--catalog", str(tmp_path / "cat.txt"), "--spots", str(sp), "--batch", "4096",
           "--ero-split", "--ero-t1", "0.58", "--ero-c0", "2", "--ero-t3", "0.59", "--ero-c1", "3",
           "--max-height", "-54", "--quiet", "--out", str(out)] + list(extra)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    return r, out


def test_spots_cli_writes_the_survivor_file(ctx, tmp_path, mst):
    fs = [f for f in mst["finds"] if f["start"]][:12]
    spots = [(f["seed"] & M48, f["cx"], f["cz"]) for f in fs]
    lines = ["# spots", "not a spot"] + ["%d %d %d" % s for s in spots]
    print("deepdive_scan --spots, %d spots" % len(spots), flush=True)
    r, out = run_cli(tmp_path, lines)
    assert r.returncode == 0, r.stderr
    _, finds, _ = ctx.dd_terrain(spots, max_height=-54)
    rows = out.read_text().splitlines()
    assert rows[0] == "seed,chunkX,chunkZ,layoutX,airHi,airLo"
    want = ["%d,%d,%d,%d,0,0" % (h[0], h[3], h[4], O.carver_seed(h[0] & M48, h[3], h[4])) for h in finds]
    assert sorted(rows[1:]) == sorted(want)
    # the confirmed huts pass the oracle's phase-1 chain at maxHeight -54, so each world seed has its row
    for f in fs:
        assert "%d,%d,%d,%d,0,0" % (f["seed"], f["cx"], f["cz"], O.carver_seed(f["seed"] & M48, f["cx"], f["cz"])) in rows
    assert ("spots: %d spots (0 not a start), %d terrain survivors" % (len(spots), len(finds))) in r.stdout


def test_spots_cli_phase2_and_non_starts(ctx, tmp_path, mst):
    """--phase2-max-y keeps the survivors lysh's hut verdict accepts; a chunk that is not a start is reported,
    skipped, and makes the exit status 1, as in the reference."""
    fs = [f for f in mst["finds"] if f["start"]][:12]
    spots = [(f["seed"] & M48, f["cx"], f["cz"]) for f in fs]
    bad = non_starts(1, 11)[0]
    r, out = run_cli(tmp_path, ["%d %d %d" % s for s in spots + [bad]], ["--phase2-max-y", "100"])
    assert r.returncode == 1
    assert "*** not a mineshaft start: S=%d chunk (%d,%d)" % bad in r.stderr
    _, finds, _ = ctx.dd_terrain(spots, max_height=-54)
    verdicts = ctx.hut_eval([(h[0], h[3], h[4]) for h in finds], whf.biome_tree("1.19.x~26.1"), max_y=100)
    keep = [h for h, v in zip(finds, verdicts) if v["ok"]]
    rows = out.read_text().splitlines()
    assert rows[0] == "seed,chunkX,chunkZ,layoutX,airHi,airLo"
    assert sorted(rows[1:]) == sorted("%d,%d,%d,%d,0,0" % (h[0], h[3], h[4], O.carver_seed(h[0] & M48, h[3], h[4]))
                                      for h in keep)
    assert "spots: %d spots (1 not a start)" % len(spots) in r.stdout
