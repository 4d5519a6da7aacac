"""CPU-side checks of the C ABI library and the phase1_scan host (no GPU compute)."""
import ctypes
import os
import re
import subprocess

import pytest

from product import whf

HEADER = os.path.join(whf.HERE, "..", "include", "whf.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|double|const char \*)\s*\*?\s*(whf_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    L = whf.load()
    names = declared_functions()
    assert names, "no declarations parsed"
    assert set(names) == set(whf.SYMBOLS)
    for n in names:
        assert hasattr(L, n), n


def test_version_flags_follow_lysh_names():
    assert whf.version_flags("1.20") == 0
    assert whf.version_flags("1.21") == 0
    assert whf.version_flags("1.19.x~26.1") == 0
    assert whf.version_flags("1.18.2") == whf.V_118
    assert whf.version_flags("1.18") == whf.V_118
    assert whf.version_flags("26.2") == whf.V_262
    assert whf.version_flags("26.3") == whf.V_262
    with pytest.raises(whf.WhfError):
        whf.version_flags("1.17")


def test_strerror_texts():
    L = whf.load()
    for code in (whf.OK, whf.EINVAL, whf.ENOMEM, whf.EOVERFLOW, whf.ENOTSUP):
        assert L.whf_strerror(code)
    assert L.whf_strerror(12345) == b"unknown error"


def test_open_rejects_null_out():
    L = whf.load()
    assert L.whf_open(None, None) == whf.EINVAL


def run_cli(*args, stdin=None):
    return subprocess.run([whf.PHASE1_SCAN, *args], input=stdin, capture_output=True, text=True, timeout=60)


@pytest.mark.skipif(not os.path.exists(whf.PHASE1_SCAN), reason="phase1_scan not built")
def test_cli_argument_errors_exit_2_before_touching_the_gpu():
    for args in (["--bogus"], ["--seed"], ["--min-region-x", "3", "--max-region-x", "3", "--seed", "1"],
                 ["--self-test", "--seed", "1"], ["--quad-bases", "x.txt"], ["--version", "1.17", "--seed", "1"],
                 ["--seed-gen", "1", "0"], ["--seed-gen", "5", "10", "--seed", "3"], ["--max-height"],
                 ["--id", "7", "--seed", "1"], ["--seed", "1", "--id"]):
        r = run_cli(*args)
        assert r.returncode == 2, (args, r.returncode, r.stderr)
        assert "usage: phase1_scan" in r.stdout
    r = run_cli()
    assert r.returncode == 2
    assert run_cli("--help").returncode == 0


def test_header_compiles_as_c99():
    """include/whf.h is plain C: it compiles on its own with a C compiler and no warnings."""
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "t.c")
        with open(src, "w") as f:
            f.write('#include "whf.h"\nint main(void) { whf_hit h; (void)h; return 0; }\n')
        r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-I", os.path.dirname(HEADER), "-c",
                            src, "-o", os.path.join(d, "t.o")], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr


@pytest.mark.skipif(not os.path.exists(whf.DEEPDIVE_SCAN), reason="deepdive_scan not built")
def test_deepdive_cli_argument_errors_exit_2_before_touching_the_gpu():
    base = ["--s-first", "1", "--count", "4", "--min-region-x", "0", "--max-region-x", "1", "--min-region-z", "0",
            "--max-region-z", "1"]
    for args in (["--bogus"], base[:4], base[:-2], base[:6] + ["0"] + base[7:], base + ["--version", "1.17"],
                 base + ["--floor"], ["--count", "x"] + base[2:]):
        r = subprocess.run([whf.DEEPDIVE_SCAN, *args], capture_output=True, text=True, timeout=60)
        assert r.returncode == 2, (args, r.returncode, r.stderr)
        assert "usage: deepdive_scan" in r.stdout
    assert subprocess.run([whf.DEEPDIVE_SCAN, "--help"], capture_output=True, timeout=60).returncode == 0


def test_deepdive_spots_cli_errors_before_touching_the_gpu(tmp_path):
    """--spots mode: bad options exit 2 and a missing spots file exits 1, all before the device is opened; the
    options of hut_lookup_v4's terrain command line that have no effect here are accepted."""
    sp = tmp_path / "spots.txt"
    sp.write_text("# no spots\n")
    for args in (["--spots"], ["--spots", str(sp), "--version", "1.17"], ["--spots", str(sp), "--max-height", "x"],
                 ["--spots", str(sp), "--phase2-max-y"], ["--spots", str(sp), "--catalog"],
                 ["--spots", str(sp), "--batch"]):
        r = subprocess.run([whf.DEEPDIVE_SCAN, *args], capture_output=True, text=True, timeout=60)
        assert r.returncode == 2, (args, r.returncode, r.stderr)
    r = subprocess.run([whf.DEEPDIVE_SCAN, "--catalog", "cat.txt", "--spots", str(tmp_path / "missing.txt"),
                        "--batch", "4096", "--ero-split", "--ero-t1", "0.5", "--ero-c0", "2", "--ero-t3", "0.6",
                        "--ero-c1", "3", "--max-height", "-54", "--quiet", "--out", str(tmp_path / "o.csv")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "cannot read" in r.stderr


def test_biome_map_cli_errors_exit_2_before_touching_the_gpu():
    """bin/biome_map takes the map tool's command line (biome_map.c:62-84): the same usage errors exit 2."""
    base = ["--seed", "5", "--x0", "0", "--z0", "0", "--w", "4", "--h", "4"]
    for args in (base[:-2], base + ["--ss", "5"], base + ["--step", "0"], ["--seed", "5", "--x0", "0", "--z0", "0",
                 "--w", "5000", "--h", "4"], base + ["--ver", "1.17"], base + ["--bogus", "1"], ["--stdin"],
                 base + ["--seed", "x"], ["--seed", "5", "--x0", str(1 << 31), "--z0", "0", "--w", "4", "--h", "4"]):
        r = subprocess.run([whf.BIOME_MAP, *args], capture_output=True, text=True, timeout=60)
        assert r.returncode == 2, (args, r.returncode, r.stderr)


@pytest.mark.skipif(not os.path.exists(whf.DEEPDIVE_SCAN), reason="deepdive_scan not built")
def test_solver_command_line_errors_exit_2_before_touching_the_gpu(tmp_path):
    """deepdive_scan --solve-only --solve5 refuses what the reference's solver refuses, with status 2 and before
    opening a device (hut_lookup_v4_source.cu:805, 841, 1297-1315)."""
    plan = tmp_path / "plan.txt"
    plan.write_text("231564844655014 22 odd 1 7.7e11 1 1\n")
    cat = tmp_path / "cat.txt"
    cat.write_text("231564844655014 331 331\n")
    base = ["--catalog", str(cat), "--solve-only", "--solve5"]
    cases = [
        ["--solve-only", "--solve5", "--plan", str(plan), "--first-index", "0", "--count", "1"],   # no --catalog
        base + ["--plan", str(plan), "--first-index", "0", "--count", "1", "--fused-blk", "100"],
        base + ["--plan", str(plan), "--first-index", "0", "--count", "1", "--fused-gpb", "0"],
        base + ["--plan", str(plan), "--first-index", "0", "--count", "1", "--fused-slots", "9"],
        base + ["--plan", str(plan), "--first-index", "0", "--count", "1", "--task-batch", "0"],
        base + ["--plan", str(plan), "--first-index", str((1 << 20) - 1), "--count", "2"],       # past the bundle
        base + ["--plan", str(plan), "--first-index", "-1", "--count", "1"],
        base + ["--plan", str(plan), "--first-index", "0", "--count", "0"],
        base + ["--first-index", "0", "--count", "1"],                                            # no plan, no regime
        base + ["--regime", "32", "--first-index", "0", "--count", "1"],
        ["--catalog", str(cat), "--solve-only", "--plan", str(plan), "--first-index", "0", "--count", "1"],
    ]
    bad_plan = tmp_path / "bad.txt"
    bad_plan.write_text("231564844655014 22 weird\n")
    cases.append(base + ["--plan", str(bad_plan), "--first-index", "0", "--count", "1"])
    bad_plan2 = tmp_path / "bad2.txt"
    bad_plan2.write_text("231564844655014 2\n")
    cases.append(base + ["--plan", str(bad_plan2), "--first-index", "0", "--count", "1"])
    for args in cases:
        r = subprocess.run([whf.DEEPDIVE_SCAN, *args], capture_output=True, text=True, timeout=60)
        assert r.returncode == 2, (args, r.returncode, r.stderr)
        assert "whf_open" not in r.stderr
    # an unreadable plan is status 1, as in the reference
    r = subprocess.run([whf.DEEPDIVE_SCAN, *base, "--plan", str(tmp_path / "absent.txt"), "--first-index", "0",
                        "--count", "1"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "cannot read plan" in r.stderr


def test_plan_parsing():
    assert whf.parse_plan("# comment\n\n231564844655014 22 odd 1 7.7e11 1 1\n156186219368766 10 u2\n1 02\n") == [
        (231564844655014, 2, 2, 1), (156186219368766, 1, 0, 2), (1, 0, 2, 1)]
    for bad in ("1 3x odd\n", "1 22 weird\n", "1 222\n"):
        with pytest.raises(ValueError):
            whf.parse_plan(bad)
