"""Whole-world Seed Check of single seeds against the reference's documented answers.

docs/HUNT54.txt:19-20 gives two seeds scanned over the whole world (hunt.py:72 WORLD, hut regions [-58594, 58594) on
both axes: 1.37e10 regions) at the -54 hunt, each with one confirmed hut: 1473884015873122059 at /tp -15792 -54 -17168
(hunt.py:86 SCAN_REFERENCE) and 1291329195661654798 at /tp 2137600 -54 26493184. Here phase 1 runs over every region
and lysh's hut verdict (phase 2, or-lower -54, the 1.19-26.1 tree) over its survivors; the finds must be exactly that
hut, in both precisions, with identical phase-1 counters."""
import numpy as np
import pytest

from product import whf

pytestmark = pytest.mark.gpu

WORLD = (-58594, 58594, -58594, 58594)
CASES = [(1473884015873122059, (-15792, -54, -17168)), (1291329195661654798, (2137600, -54, 26493184))]


@pytest.mark.parametrize("seed,hut", CASES)
def test_whole_world_known_hut(seed, hut):
    vf = whf.version_flags("1.20")
    tree = whf.biome_tree("1.20")
    got = {}
    with whf.Context(0) as ctx:
        for prec in (whf.PREC_FP32_GATE, whf.PREC_FP64):
            print("seed %d precision %d" % (seed, prec), flush=True)
            hits, st = ctx.scan(WORLD, seeds=np.array([seed], dtype=np.int64), max_height=-50, vflags=vf,
                                capacity=1 << 12, precision=prec, phase2=(whf.P2_OR_LOWER, -54, tree))
            assert st["checks"] == (WORLD[1] - WORLD[0]) * (WORLD[3] - WORLD[2])
            finds = [(int(h[3]) * 16, int(h[5]), int(h[4]) * 16) for h in hits]
            assert finds == [hut], (prec, finds)
            got[prec] = (st["funnel"], st["tiers"], st["accepted"], st["phase2"])
    assert got[whf.PREC_FP32_GATE] == got[whf.PREC_FP64]
