#!/usr/bin/env python3
"""Regenerates the golden fixtures under tests/golden/ from the reference CPU scanner.

Needs the reference checker built by `make -C oracle ref` (oracle/_ref, compiled from the reference's own
lysh sources) and the reference tree for its headers, so it runs in the development container only. The
fixtures it writes are plain data (inputs and the reference's outputs) and are what the tests read.

    python tests/golden/make_golden.py
"""
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = os.environ.get("WHF_REFERENCE", "/root/reference")
LYSH_SRC = os.path.join(REF, "engine/tools/lysh/src")
REF_BUILD = os.path.join(REPO, "oracle", "_ref")


def run(cmd, **kw):
    return subprocess.run(cmd, check=True, capture_output=True, text=True, **kw).stdout


def main():
    tool = os.path.join(REF_BUILD, "gen_golden")
    run(["gcc", "-O2", "-std=gnu11", "-ffp-contract=off", "-w", "-I" + LYSH_SRC,
         os.path.join(HERE, "gen_golden.c"), os.path.join(REF_BUILD, "liblysh.a"), "-lm", "-lpthread", "-o", tool])
    vec = json.loads(run([tool]))
    with open(os.path.join(HERE, "lysh_vectors.json"), "w") as f:
        json.dump(vec, f, separators=(",", ":"))

    bench = os.path.join(REF_BUILD, "lysh_cpu_bench")
    cases = [
        # (name, stream, count, rect, version)
        ("c1_v120", 1, 65536, (-2, 2, -2, 2), "1.20"),
        ("c1_v118", 1, 16384, (-2, 2, -2, 2), "1.18"),
        ("c1_v262", 1, 16384, (-2, 2, -2, 2), "26.2"),
        ("c2_prefix_v120", 2, 16384, (-8, 8, -8, 8), "1.20"),
        ("c2_prefix_v121", 2, 4096, (-8, 8, -8, 8), "1.21"),
    ]
    funnels = {}
    for name, stream, count, (a, b, c, d), ver in cases:
        out = run([bench, "--stream", str(stream), "--count", str(count), "--rx0", str(a), "--rx1", str(b),
                   "--rz0", str(c), "--rz1", str(d), "--version", ver, "--max-y", "-50", "--threads", "8"])
        r = json.loads(out)
        funnels[name] = dict(stream=stream, count=count, rect=[a, b, c, d], version=ver, max_y=-50,
                             checks=r["checks"], accepted=r["accepted"], reached_ladder=r["reached_ladder"],
                             funnel=r["funnel"], tiers=r["tiers"])
    # the reference's known answer (hunt.py:86, tools/test_hunt.py:264-271): lysh scan --list over +-79 regions
    lysh = os.path.join(REF_BUILD, "lysh")
    out = run([lysh, "scan", "--seed", "1473884015873122059", "--rx0", "-79", "--rx1", "79", "--rz0", "-79",
               "--rz1", "79", "--max-y", "-50", "--phase2-max-y", "-50", "--version", "1.19.x~26.1", "--threads", "8",
               "--list", "--quiet"])
    hits = [list(map(int, l.split()[1:])) for l in out.splitlines() if l.startswith("hit ")]
    funnels["known_answer"] = dict(seed=1473884015873122059, rect=[-79, 79, -79, 79], version="1.19.x~26.1",
                                   max_y=-50, hits=hits)
    with open(os.path.join(HERE, "lysh_funnels.json"), "w") as f:
        json.dump(funnels, f, indent=1)
    print("wrote", os.path.join(HERE, "lysh_vectors.json"), "and lysh_funnels.json")


if __name__ == "__main__":
    sys.exit(main())
