"""hunt.py's machine helpers for an AMD GPU (low-witch-finder_amd/amd_gpu.py): the status fields parsed from amd-smi's
JSON (fixtures tests/golden/amd_smi_metric.json and amd_smi_static.json, captured from ROCm 7.2's amd-smi on an
MI355X and cut to the fields read), and a live call where a GPU is present."""
import json
import os
import shutil

import importlib.util

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
_spec = importlib.util.spec_from_file_location("amd_gpu", os.path.join(HERE, "..", "low-witch-finder_amd", "amd_gpu.py"))
amd_gpu = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(amd_gpu)

GOLD = os.path.join(HERE, "golden")


def load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


def test_parse_captured():
    s = amd_gpu.parse_stats(load("amd_smi_metric.json"), load("amd_smi_static.json"))
    assert s == dict(name="AMD Instinct MI355 OAM", temp=48, util=3, mem_used=0, mem_total=294896, watts=271.0,
                     arch="gfx950")


def test_parse_missing_fields():
    m = load("amd_smi_metric.json")
    m["gpu_data"][0]["power"]["socket_power"] = "N/A"
    del m["gpu_data"][0]["temperature"]["hotspot"]
    s = amd_gpu.parse_stats(m, None)
    assert s["watts"] is None and s["temp"] is None and s["name"] is None and s["mem_total"] == 294896
    assert s["arch"] is None
    assert amd_gpu.parse_stats({}, None) is None


def test_preflight_rows(tmp_path):
    """The GPU mode's checklist (hunt.py:468-519) on an AMD card: card, architecture, free memory, tools, and the
    build hint that names ROCm's compiler when it is missing."""
    g = amd_gpu.parse_stats(load("amd_smi_metric.json"), load("amd_smi_static.json"))
    bin_dir = tmp_path / "bin"
    bin_dir.mkdir()
    rows = amd_gpu.preflight_rows(-58, str(bin_dir), stats=g, hipcc="")
    assert rows[0] == (True, "AMD Instinct MI355 OAM", None)
    assert rows[1][0] and "gfx950" in rows[1][1]
    assert rows[2][0] and "video memory free" in rows[2][1]
    assert rows[3][0] is False and rows[3][2] == "gpu-build" and "hipcc" in rows[3][1]
    rows = amd_gpu.preflight_rows(-58, str(bin_dir), stats=g, hipcc="/opt/rocm/bin/hipcc")
    assert rows[3][2] == "gpu-build" and "build_rocm.sh" in rows[3][1] and "hut_lookup_v4" in rows[3][1]
    for n in ("phase1_scan", "lysh"):
        (bin_dir / n).write_text("")
    rows = amd_gpu.preflight_rows(-52, str(bin_dir), stats=g, hipcc="")
    assert rows[-1] == (True, "tools built", None)
    rows = amd_gpu.preflight_rows(-52, str(bin_dir), stats={}, hipcc="")
    assert rows[0][0] is False and "no AMD GPU" in rows[0][1]
    other = dict(g, arch="gfx942", mem_used=g["mem_total"] - 100)
    rows = amd_gpu.preflight_rows(-52, str(bin_dir), stats=other, hipcc="")
    assert rows[1][0] is False and "gfx942" in rows[1][1]
    assert rows[2][0] is False and "only 100 MiB" in rows[2][1]


def test_build_script_links_the_programs_under_hunt_names(tmp_path):
    """build_rocm.sh links phase1_scan, deepdive_scan (as hut_lookup_v4) and biome_map into a checkout's bin/, and the
    links run (the programs find libwhf_hip.so through their run path)."""
    import subprocess
    script = os.path.join(HERE, "..", "low-witch-finder_amd", "build_rocm.sh")
    prog = os.path.join(HERE, "..", "low-witch-finder_amd", "bin", "deepdive_scan")
    if not os.path.exists(prog) or not (shutil.which("hipcc") or os.path.exists("/opt/rocm/bin/hipcc")):
        pytest.skip("no hipcc or programs not built")
    r = subprocess.run(["bash", script, str(tmp_path)], capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stdout + r.stderr
    for n in ("phase1_scan", "hut_lookup_v4", "biome_map"):
        assert os.path.islink(tmp_path / "bin" / n)
    r = subprocess.run([str(tmp_path / "bin" / "hut_lookup_v4"), "--help"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "--solve5" in r.stdout


def test_processes_counts_nothing_here():
    assert amd_gpu.gpu_processes() >= 0


@pytest.mark.gpu
def test_live_stats():
    if not shutil.which("amd-smi"):
        pytest.skip("amd-smi not on PATH")
    s = amd_gpu.gpu_stats()
    assert s is not None and s["mem_total"] and s["mem_total"] > 0, s
