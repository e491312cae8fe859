#!/bin/bash
# Compiles the reference's CPU sources for the symbolic phase, the certificate and the CPU
# simplex route (Phase I / Phase II over the sparse LU), where they lie under /root/reference,
# together with oracle/ref_shim.cpp, into oracle/_ref/libref.so (-O2 with the reference's IEEE
# flags: no fast-math).  The CPU tests load it as a cross-check of this repository's
# restatements and bench.py times its simplex route as the CPU baseline; nothing is copied from
# the reference and nothing else is built.
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
REF="${REFERENCE_ROOT:-/root/reference}"
OUT="$HERE/_ref"
[ -f "$REF/src/ipm/symbolic.cpp" ] || { echo "build_ref: no reference sources at $REF" >&2; exit 0; }
mkdir -p "$OUT"
g++ -std=c++17 -O2 -fPIC -shared -fno-fast-math -w \
    -I"$REF/src" -I"$REF/include" \
    "$REF/src/ipm/symbolic.cpp" "$REF/src/verifier/certificate.cpp" \
    "$REF/src/core/problem.cpp" "$REF/src/core/sparse_matrix.cpp" \
    "$REF/src/simplex/phase1.cpp" "$REF/src/simplex/primal.cpp" \
    "$REF/src/numerics/sparse_lu.cpp" \
    "$HERE/ref_shim.cpp" -o "$OUT/libref.so.tmp"
mv "$OUT/libref.so.tmp" "$OUT/libref.so"
