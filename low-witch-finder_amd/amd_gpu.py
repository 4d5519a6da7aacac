"""The machine-side helpers hunt.py keeps for an NVIDIA card, for an AMD Instinct GPU: what its status line and its
GPU-process count read (hunt.py:151-171 gpu_stats, hunt.py:229-250 gpu_processes), from ROCm's amd-smi and /proc,
and the preflight checklist of a GPU mode (hunt.py:468-519 preflight_rows) for this engine.

    stats = gpu_stats()            # {name, temp, util, mem_used, mem_total, watts, arch} of GPU 0, or None
    n = gpu_processes()            # running phase1_scan / deepdive_scan (also linked as hut_lookup_v4) processes
    rows = preflight_rows(-58, bin_dir)   # [(ok, text, fix)]

A field the tool does not report ("N/A") is None, as hunt.py's own parser does for nvidia-smi.
"""
import json
import os
import shutil
import subprocess

ENGINE_NAMES = ("phase1_scan", "deepdive_scan", "hut_lookup_v4")
ENGINE_ARCH = "gfx950"   # the engine's only code object target


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
        arch = s.get("target_graphics_version")
    except (KeyError, IndexError, TypeError):
        name = arch = None
    temp = m.get("temperature", {})
    t = _value(temp.get("edge"))
    if t is None:   # Instinct boards report the hotspot, not an edge sensor
        t = _value(temp.get("hotspot"))
    mem = m.get("mem_usage", {})
    return dict(name=name, temp=t, util=_value(m.get("usage", {}).get("gfx_activity")),
                mem_used=_value(mem.get("used_vram")), mem_total=_value(mem.get("total_vram")),
                watts=_value(m.get("power", {}).get("socket_power"), float), arch=arch)


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


def preflight_rows(level, bin_dir, stats=None, min_free_vram=3000, hipcc=None, tool_names=None):
    """hunt.py's preflight checklist of a GPU mode (hunt.py:468-519) on an AMD card: [(ok, text, fix)], fix None or
    "build" / "gpu-build" / "close-minecraft" as hunt.py's buttons take them. `stats` is gpu_stats() unless given;
    `hipcc` the compiler's path (looked up unless given); `tool_names` overrides which tools in bin_dir are needed.
    The Minecraft, Java, memory and disk rows are hunt.py's own and stay there."""
    rows = []

    def add(ok, good, bad, fix=None):
        rows.append((bool(ok), good if ok else bad, None if ok else fix))
    g = gpu_stats() if stats is None else stats
    add(g, (g or {}).get("name") or "AMD GPU",
        "no AMD GPU found (amd-smi): pick a CPU mode (it covers -50 to -54), or hunt on a machine with one")
    if g:
        arch = g.get("arch")
        add(arch in (None, ENGINE_ARCH), "%s (%s)" % (g.get("name") or "AMD GPU", arch or ENGINE_ARCH),
            "this engine is built for %s (MI355X); the card reports %s" % (ENGINE_ARCH, arch))
        total, used = g.get("mem_total"), g.get("mem_used")
        free = None if total is None else total - (used or 0)
        add(free is None or free >= min_free_vram, "%.1f GB video memory free" % ((free or 0) / 1024.0),
            "only %s MiB of video memory free (needs %d): close programs that use the GPU" % (free, min_free_vram))
    solver = level <= -55
    names = tool_names or (("hut_lookup_v4", "mstree", "lysh", "msroom", "skyscan", "biome_at", "terrain_dump")
                           if solver else ("phase1_scan", "lysh"))
    miss = [n for n in names if not os.path.isfile(os.path.join(bin_dir, n))]
    if miss:
        gpu_part = [n for n in miss if n in ("hut_lookup_v4", "phase1_scan")]
        cc = hipcc if hipcc is not None else (shutil.which("hipcc") or
                                                 ("/opt/rocm/bin/hipcc" if os.path.isfile("/opt/rocm/bin/hipcc") else ""))
        if gpu_part and not cc:
            add(False, "", "the GPU engine isn't built, and building it needs ROCm's HIP compiler (hipcc, in the "
                           "rocm-hip-sdk package): install ROCm, then run low-witch-finder_amd/build_rocm.sh", "gpu-build")
        elif gpu_part:
            add(False, "", "the GPU engine isn't built yet: run low-witch-finder_amd/build_rocm.sh with this folder "
                           "(%s)" % ", ".join(gpu_part), "gpu-build")
        else:
            add(False, "", "tools not built yet: %s" % ", ".join(miss), "build")
    else:
        add(True, "tools built", "")
    return rows
