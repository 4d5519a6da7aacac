#!/usr/bin/env bash
# Builds the MI355X engine (libwhf_hip.so and its programs, hipcc for gfx950) and, given a Low Witch Finder
# checkout, links the programs into its bin/ under the names hunt.py runs: phase1_scan (the Seed Check GPU
# worker), hut_lookup_v4 (the Deep Dive solver and terrain stage, deepdive_scan) and biome_map (the map tab).
# It stands in for the reference's `build.sh gpu` (build.sh:51-62), which builds the CUDA engine with nvcc.
#
#   low-witch-finder_amd/build_rocm.sh [CHECKOUT]
#
# The links point back into this folder, where the programs find libwhf_hip.so (their run path is $ORIGIN/..).
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
HIPCC="${HIPCC:-$(command -v hipcc || echo /opt/rocm/bin/hipcc)}"
if [[ ! -x "$HIPCC" ]]; then
    echo "hipcc (ROCm) isn't installed, so the GPU engine can't be built here."
    exit 1
fi
echo "building the GPU engine with hipcc for gfx950 ..."
make -C "$HERE" -j"${JOBS:-8}" HIPCC="$HIPCC"
if [[ $# -ge 1 ]]; then
    BIN="$1/bin"
    mkdir -p "$BIN"
    ln -sfn "$HERE/bin/phase1_scan" "$BIN/phase1_scan"
    ln -sfn "$HERE/bin/deepdive_scan" "$BIN/hut_lookup_v4"
    ln -sfn "$HERE/bin/biome_map" "$BIN/biome_map"
    echo "linked into $BIN: phase1_scan hut_lookup_v4 biome_map"
fi
echo "done"
