"""The machine-side helpers hunt.py keeps for an NVIDIA card, for an AMD Instinct GPU: what its status line and its
GPU-process count read (hunt.py:151-171 gpu_stats, hunt.py:229-250 gpu_processes), from ROCm's amd-smi and /proc.

    stats = gpu_stats()            # {name, temp, util, mem_used, mem_total, watts} of GPU 0, or None
    n = gpu_processes()            # running phase1_scan / deepdive_scan processes

A field the tool does not report ("N/A") is None, as hunt.py's own parser does for nvidia-smi.
"""
import json
import os
import shutil
import subprocess

ENGINE_NAMES = ("phase1_scan", "deepdive_scan")


def _value(field, cast=int):
    if isinstance(field, dict):
        field = field.get("value")
    try:
        return cast(float(field))
    except (TypeError, ValueError):
        return None


def parse_stats(metric, static):
    """hunt.py's stats dict from `amd-smi metric -u -t -m -p --json` and `amd-smi static --asic --json` (parsed)."""
    try:
        m = metric["gpu_data"][0] if isinstance(metric, dict) else metric[0]
    except (KeyError, IndexError, TypeError):
        return None
    try:
        s = (static["gpu_data"][0] if isinstance(static, dict) else static[0])["asic"]
        name = s.get("market_name")
    except (KeyError, IndexError, TypeError):
        name = None
    temp = m.get("temperature", {})
    t = _value(temp.get("edge"))
    if t is None:   # Instinct boards report the hotspot, not an edge sensor
        t = _value(temp.get("hotspot"))
    mem = m.get("mem_usage", {})
    return dict(name=name, temp=t, util=_value(m.get("usage", {}).get("gfx_activity")),
                mem_used=_value(mem.get("used_vram")), mem_total=_value(mem.get("total_vram")),
                watts=_value(m.get("power", {}).get("socket_power"), float))


def gpu_stats(device=0):
    smi = shutil.which("amd-smi")
    if not smi:
        return None
    try:
        metric = subprocess.run([smi, "metric", "-g", str(device), "-u", "-t", "-m", "-p", "--json"],
                                capture_output=True, text=True, timeout=20)
        static = subprocess.run([smi, "static", "-g", str(device), "--asic", "--json"], capture_output=True, text=True,
                                timeout=20)
        if metric.returncode != 0:
            return None
        return parse_stats(json.loads(metric.stdout), json.loads(static.stdout) if static.returncode == 0 else None)
    except (OSError, subprocess.SubprocessError, ValueError):
        return None


def gpu_processes():
    n = 0
    if os.path.isdir("/proc"):
        for pid in os.listdir("/proc"):
            if pid.isdigit():
                try:
                    with open("/proc/%s/comm" % pid, encoding="utf-8") as f:
                        if f.read().startswith(ENGINE_NAMES):
                            n += 1
                except OSError:
                    pass
    return n
