#!/usr/bin/env python3
"""Writes tests/golden/mstree_vectors.json: Deep Dive vectors from the reference's own mineshaft tool.

`mstree` (engine/tools/mstree/mstree.cpp, built from the reference's sources by `make -C oracle ref` into
oracle/_ref/mstree) prints, for a world seed and a chunk, whether the chunk holds a mineshaft start (exit 3 when
it does not), the room floor, and the layout's pieces (kind x0 y0 z0 x1 y1 z1 facing parent rails). This script
records, for fixed inputs:
  random  - random (seed, chunk) pairs: the start flag, and for starts the floor, piece count and the SHA-256 of
            the piece lines;
  starts  - start chunks (picked with the oracle's start test, confirmed by mstree): the same fields;
  layouts - three small layouts in full;
  finds   - the confirmed deep huts of the reference's data/finds.tsv whose coordinates are inside the hut:
            their seed, coordinates and depth, and mstree's room floor for the hut chunk.
Run in the development container (it needs /root/reference and the oracle build); the tests read only the JSON.
"""
import hashlib
import json
import os
import random
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_ctypes as O  # noqa: E402

MSTREE = os.path.join(REPO, "oracle", "_ref", "mstree")
FINDS = "/root/reference/data/finds.tsv"
M48 = (1 << 48) - 1


def mstree(seed, cx, cz):
    p = subprocess.run([MSTREE, str(seed), str(cx), str(cz)], capture_output=True, text=True, timeout=60)
    if p.returncode == 3:
        return dict(start=False)
    assert p.returncode == 0, p.stderr
    lines = p.stdout.strip().split("\n") if p.stdout.strip() else []
    floor = int(p.stderr.strip().split()[-1])
    return dict(start=True, room_floor=floor, pieces=len(lines),
                sha256=hashlib.sha256("\n".join(lines).encode()).hexdigest(), lines=lines)


def main():
    rng = random.Random(20260312)
    out = dict(tool="engine/tools/mstree/mstree.cpp", random=[], starts=[], layouts=[], finds=[])
    for _ in range(300):
        seed = rng.getrandbits(64) - (1 << 63)
        cx, cz = rng.randint(-1875000, 1874999), rng.randint(-1875000, 1874999)
        r = mstree(seed, cx, cz)
        r.pop("lines", None)
        out["random"].append(dict(seed=seed, cx=cx, cz=cz, **r))
    small = []
    while len(out["starts"]) < 200:
        seed = rng.getrandbits(64) - (1 << 63)
        cx, cz = rng.randint(-1875000, 1874999), rng.randint(-1875000, 1874999)
        if not O.is_start(O.carver_seed(seed & M48, cx, cz)):
            continue
        r = mstree(seed, cx, cz)
        assert r["start"], (seed, cx, cz)
        if r["pieces"] <= 24 and len(small) < 3:
            small.append(dict(seed=seed, cx=cx, cz=cz, room_floor=r["room_floor"], lines=r["lines"]))
        r.pop("lines")
        out["starts"].append(dict(seed=seed, cx=cx, cz=cz, **r))
    out["layouts"] = small
    for line in open(FINDS):
        if line.startswith("#"):
            continue
        f = line.rstrip("\n").split("\t")
        if len(f) < 5 or not f[2]:
            continue
        depth, seed, x, y, z = int(f[0]), int(f[1]), int(f[2]), int(f[3]), int(f[4])
        r = mstree(seed, x >> 4, z >> 4)
        out["finds"].append(dict(depth=depth, seed=seed, x=x, y=y, z=z, cx=x >> 4, cz=z >> 4, start=r["start"],
                                 room_floor=r.get("room_floor"), pieces=r.get("pieces"), sha256=r.get("sha256")))
    with open(os.path.join(HERE, "mstree_vectors.json"), "w") as f:
        json.dump(out, f, indent=0)
    print("random starts:", sum(v["start"] for v in out["random"]), "layouts:", len(small), "finds:", len(out["finds"]))


if __name__ == "__main__":
    main()
