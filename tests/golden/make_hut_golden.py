#!/usr/bin/env python3
"""Regenerates tests/golden/hut_vectors.json from the reference CPU scanner's carvers and hut evaluation.

Needs oracle/_ref/liblysh.a (`make -C oracle ref`) and the reference tree for its sources, so it runs in the
development container only. gen_hut_golden.c documents the sections; the fixture is data only.

    python tests/golden/make_hut_golden.py
"""
import json
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = os.environ.get("WHF_REFERENCE", "/root/reference")
LYSH_SRC = os.path.join(REF, "engine/tools/lysh/src")
REF_BUILD = os.path.join(REPO, "oracle", "_ref")


def main():
    tool = os.path.join(REF_BUILD, "gen_hut_golden")
    subprocess.run(["gcc", "-O2", "-std=gnu11", "-ffp-contract=off", "-w", "-I" + LYSH_SRC,
                    os.path.join(HERE, "gen_hut_golden.c"), os.path.join(REF_BUILD, "liblysh.a"), "-lm",
                    "-lpthread", "-o", tool], check=True)
    out = subprocess.run([tool], check=True, capture_output=True, text=True).stdout
    vec = json.loads(out)
    with open(os.path.join(HERE, "hut_vectors.json"), "w") as f:
        json.dump(vec, f, separators=(",", ":"))
    print({k: len(v) for k, v in vec.items()})


if __name__ == "__main__":
    main()
