"""Launch shapes change only the speed: every shape the autotuner can choose gives the default's exact answer, and
whf_open applies the device's tuned.json entry (low-witch-finder_amd/tune.py writes it)."""
import json
import os

import pytest

from product import whf

pytestmark = pytest.mark.gpu

SHAPES = [dict(ero_wg=256, ero_ilp=1), dict(ero_wg=512, ero_ilp=1), dict(ero_wg=256, ero_ilp=2, ero_wg_per_cu=8),
          dict(ero_ilp=1, ero_wg_per_cu=3), dict(gate_wg_per_cu=1), dict(gate_wg_per_cu=32)]


@pytest.mark.parametrize("version,prec", [("1.20", whf.PREC_FP32_GATE), ("1.18", whf.PREC_FP64),
                                          ("26.2", whf.PREC_FP32_GATE)])
def test_every_shape_gives_the_default_answer(version, prec):
    vf = whf.version_flags(version)
    kw = dict(splitmix=(5, 3000), max_height=-30, vflags=vf, precision=prec, capacity=1 << 16)
    rect = (-12, 12, -12, 12)
    with whf.Context(0) as ctx:
        print("default shape, %s" % version, flush=True)
        ref = ctx.scan(rect, **kw)
        assert ref[1]["accepted"] > 0
        for shape in SHAPES:
            ctx.set_tuning(vflags=vf, **shape)
            print("shape %r" % shape, flush=True)
            got = ctx.scan(rect, **kw)
            assert got[0] == ref[0], shape
            for k in ("funnel", "tiers", "accepted", "reached_ladder", "checks"):
                assert got[1][k] == ref[1][k], (shape, k)
        # a narrow rect takes the 8-seed tiles
        ctx.set_tuning(ero_wg=256, ero_ilp=2, ero_wg_per_cu=-1, gate_wg_per_cu=8)
        narrow = ctx.scan((0, 3, -1, 1), **kw)
        ctx.set_tuning(ero_wg=512, ero_ilp=1, ero_wg_per_cu=4)
        assert ctx.scan((0, 3, -1, 1), **kw)[0] == narrow[0]


def test_whf_open_applies_the_tuned_file(tmp_path, monkeypatch):
    with whf.Context(0) as ctx:
        device = ctx.device_name.split(" (")[0]
    path = tmp_path / "tuned.json"
    path.write_text(json.dumps({device: {"build": whf.build_id(), "ero_ilp": 1, "gate_wg_per_cu": 4,
                                         "versions": {"1.18": {"ero_wg": 512}}}}))
    monkeypatch.setenv("WHF_TUNED", str(path))
    with whf.Context(0) as ctx:
        t0 = ctx.tuning(0)
        t1 = ctx.tuning(whf.V_118)
    assert (t0["ero_ilp"], t0["gate_wg_per_cu"], t0["ero_wg"]) == (1, 4, 256)
    assert (t1["ero_ilp"], t1["gate_wg_per_cu"], t1["ero_wg"]) == (1, 4, 512)
    # an entry of another build is ignored
    path.write_text(json.dumps({device: {"build": "x", "ero_ilp": 1}}))
    with whf.Context(0) as ctx:
        assert ctx.tuning(0)["ero_ilp"] == 2
