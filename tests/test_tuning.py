"""The tuned-configuration file as whf_open reads it (whf_load_tuning, no device needed): the device's entry or the
"default" entry, per-version overrides, and the rejection of entries measured with another build or out of range."""
import json

from product import whf

DEFAULTS = dict(ero_wg=256, ero_ilp=2, ero_wg_per_cu=-1, gate_wg_per_cu=16, chunk_checks=1 << 24)


def write(tmp_path, doc):
    p = tmp_path / "tuned.json"
    p.write_text(json.dumps(doc) if not isinstance(doc, str) else doc)
    return str(p)


def test_device_entry_and_version_overrides(tmp_path):
    build = whf.build_id()
    path = write(tmp_path, {
        "AMD Instinct MI355X": {"build": build, "ero_wg": 512, "ero_ilp": 1, "gate_wg_per_cu": 4,
                                "chunk_checks": 1 << 22, "measured": {"ms": [1.0, 2.0], "note": "x"},
                                "versions": {"1.18": {"gate_wg_per_cu": 16}, "26.2": {"ero_wg": 256, "ero_ilp": 2}}},
        "default": {"build": build, "gate_wg_per_cu": 2},
    })
    ok, t = whf.load_tuning(path, "AMD Instinct MI355X", 0)
    assert ok and t == dict(ero_wg=512, ero_ilp=1, ero_wg_per_cu=-1, gate_wg_per_cu=4, chunk_checks=1 << 22)
    # 1.18 rules (and its presets) take the override, the rest of the device entry stays
    for vf in (whf.V_118, whf.V_118 | whf.V_LARGE, whf.V_118 | whf.V_SINGLE):
        ok, t = whf.load_tuning(path, "AMD Instinct MI355X", vf)
        assert ok and t["gate_wg_per_cu"] == 16 and t["ero_wg"] == 512, (vf, t)
    ok, t = whf.load_tuning(path, "AMD Instinct MI355X", whf.V_262)
    assert ok and (t["ero_wg"], t["ero_ilp"], t["gate_wg_per_cu"]) == (256, 2, 4)
    # another device falls back to the default entry
    ok, t = whf.load_tuning(path, "Some Other GPU", 0)
    assert ok and t == dict(DEFAULTS, gate_wg_per_cu=2)


def test_entries_of_another_build_or_out_of_range_are_ignored(tmp_path):
    path = write(tmp_path, {"GPU": {"build": "not-this-build", "ero_wg": 512}})
    assert whf.load_tuning(path, "GPU", 0) == (False, DEFAULTS)
    path = write(tmp_path, {"GPU": {"build": whf.build_id(), "ero_wg": 300}})
    assert whf.load_tuning(path, "GPU", 0) == (False, DEFAULTS)
    path = write(tmp_path, {"GPU": {"build": whf.build_id(), "gate_wg_per_cu": 99}})
    assert whf.load_tuning(path, "GPU", 0) == (False, DEFAULTS)


def test_missing_or_malformed_file_gives_defaults(tmp_path):
    assert whf.load_tuning(str(tmp_path / "absent.json"), "GPU", 0) == (False, DEFAULTS)
    assert whf.load_tuning(write(tmp_path, '{"GPU": {"build": '), "GPU", 0) == (False, DEFAULTS)
    assert whf.load_tuning(write(tmp_path, "[1, 2, 3]"), "GPU", 0) == (False, DEFAULTS)
