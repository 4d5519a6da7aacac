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
    assert s == dict(name="AMD Instinct MI355 OAM", temp=48, util=3, mem_used=0, mem_total=294896, watts=271.0)


def test_parse_missing_fields():
    m = load("amd_smi_metric.json")
    m["gpu_data"][0]["power"]["socket_power"] = "N/A"
    del m["gpu_data"][0]["temperature"]["hotspot"]
    s = amd_gpu.parse_stats(m, None)
    assert s["watts"] is None and s["temp"] is None and s["name"] is None and s["mem_total"] == 294896
    assert amd_gpu.parse_stats({}, None) is None


def test_processes_counts_nothing_here():
    assert amd_gpu.gpu_processes() >= 0


@pytest.mark.gpu
def test_live_stats():
    if not shutil.which("amd-smi"):
        pytest.skip("amd-smi not on PATH")
    s = amd_gpu.gpu_stats()
    assert s is not None and s["mem_total"] and s["mem_total"] > 0, s
