"""bin/biome_map (the map tool's biome ids on the GPU) against the reference's own biome_map, compiled from its
sources into oracle/_ref (engine/tools/mineshaft/biome_map.c over lysh's biome check).

Bar: byte-identical output for the same command line: the map mode over several versions, cell sizes, sample
counts and heights, and the --stdin mode over the chunks around a swamp, where the six-height test disagrees on
the swamp's edge and the real surface decides.
"""
import os
import subprocess

import pytest

from product import whf

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_MAP = os.path.join(REPO, "oracle", "_ref", "biome_map")
KNOWN = 1473884015873122059   # hut at block (-15792, -17168): chunk (-987, -1073), a swamp


def both(args, stdin=None):
    print("biome_map " + " ".join(args), flush=True)
    ours = subprocess.run([whf.BIOME_MAP, *args], input=stdin, capture_output=True, timeout=300)
    ref = subprocess.run([REF_MAP, *args], input=stdin, capture_output=True, timeout=300)
    assert ours.returncode == 0, ours.stderr
    assert ref.returncode == 0, ref.stderr
    return ours.stdout, ref.stdout


@pytest.mark.parametrize("args", [
    ["--seed", str(KNOWN), "--x0", "-18000", "--z0", "-19200", "--w", "48", "--h", "32", "--step", "96", "--ver", "1.20.1"],
    ["--seed", str(KNOWN), "--x0", "-24000", "--z0", "-24000", "--w", "32", "--h", "16", "--step", "512", "--ss", "3",
     "--ver", "26.3"],
    ["--seed", "-4172144997902289642", "--x0", "100", "--z0", "-2000", "--w", "64", "--h", "16", "--step", "16",
     "--ss", "2", "--ver", "1.18.2", "--y", "70"],
    ["--seed", "42", "--x0", "-9000", "--z0", "5000", "--w", "16", "--h", "8", "--step", "1000", "--ss", "4",
     "--ver", "1.21.4"],
    ["--seed", "7", "--x0", "-37", "--z0", "-5", "--w", "40", "--h", "12", "--step", "5", "--ver", "1.19.3"],
])
def test_map_mode_matches_the_reference(args):
    ours, ref = both(args)
    w, h = int(args[args.index("--w") + 1]), int(args[args.index("--h") + 1])
    assert len(ref) == w * h
    assert ours == ref


def test_map_mode_shows_swamps():
    """A map around the known hut has swamp cells (id 6), so the comparison above covers the swamp rule."""
    ours, _ = both(["--seed", str(KNOWN), "--x0", "-16400", "--z0", "-17800", "--w", "40", "--h", "40", "--step", "32",
                    "--ss", "2", "--ver", "1.20.1"])
    assert ours.count(bytes([6])) > 10 and len(set(ours)) > 1


@pytest.mark.parametrize("ver", ["1.20.1", "26.3"])
def test_stdin_mode_matches_the_reference(ver):
    """Chunks on swamp edges, where the six heights disagree and the real surface decides, plus others where they
    agree. Such chunks are rare (none among the 200 x 200 chunks around the known hut); the reference's own map at
    --y 40 and --y 80 differs in 113 of these 80 x 24 chunks, so they are found here."""
    area = [(cx, cz) for cz in range(-1592, -1568) for cx in range(-1260, -1180)]
    with whf.Context(0) as ctx:
        _, win = ctx.biome([(KNOWN, cx * 16, cz * 16, 64) for cx, cz in area], whf.biome_tree(ver), winners=True)
    cols = [(y >> 2) - whf.BIOME_QY0 for y in (40, 48, 56, 64, 72, 80)]
    votes = [sum(int(row[c]) == 6 for c in cols) for row in win]
    edge = [c for c, v in zip(area, votes) if 0 < v < 6]
    print("%d chunks on a swamp edge" % len(edge), flush=True)
    assert len(edge) >= 1
    chunks = edge[:64] + [c for c, v in zip(area, votes) if v == 6][:32] + area[::97]
    chunks += [(0, 0), (-987, -1073), (123456, -654321)]
    text = "".join("%d %d\n" % c for c in chunks).encode()
    ours, ref = both(["--seed", str(KNOWN), "--stdin", "--ver", ver], stdin=text)
    assert len(ref) == len(chunks)
    assert ours == ref
    assert ours[chunks.index((-987, -1073))] == 6
