#!/bin/bash
# Builds oracle/_ref/libref_geom.so from the reference's host-only C++ sources (no CUDA:
# the __CUDACC__ sections are skipped; __host__/__device__ qualifiers on two plain
# functions are defined empty). Only runs where the reference checkout exists.
set -e
REF=${FASES_REFERENCE:-/root/reference}
HERE="$(cd "$(dirname "$0")" && pwd)"
if [ ! -f "$REF/3D-solver/include/parsing3D.h" ]; then
    echo "reference sources not present: skipping oracle/_ref"
    exit 0
fi
mkdir -p "$HERE/_ref"
g++ -std=c++17 -O2 -fPIC -shared -D__host__= -D__device__= \
    -I"$REF/3D-solver/include" -I"$REF/common" \
    "$HERE/ref_geom_shim.cpp" -o "$HERE/_ref/libref_geom.so"
echo "built $HERE/_ref/libref_geom.so"
