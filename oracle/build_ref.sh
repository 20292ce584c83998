#!/bin/bash
# Builds oracle/_ref/libref.so from the reference's own headers, compiled where
# they lie under /root/reference/src, plus oracle/ref_harness.cpp.  Test
# infrastructure only; the output stays out of git.  Without the reference
# checkout this script does nothing and exits 0.
set -e
HERE="$(cd "$(dirname "$0")" && pwd)"
REF="${RTN_REFERENCE_SRC:-/root/reference/src}"
if [ ! -d "$REF" ]; then
  echo "build_ref: $REF not present; skipping reference build"
  exit 0
fi
mkdir -p "$HERE/_ref"
# -ffp-contract=off and no -march: plain SSE2 scalar FP32, no FMA, host 1/sqrtf.
g++ -std=c++17 -O2 -ffp-contract=off -fno-fast-math -fPIC -shared -w \
    -I"$REF" "$HERE/ref_harness.cpp" -o "$HERE/_ref/libref.so.tmp"
mv "$HERE/_ref/libref.so.tmp" "$HERE/_ref/libref.so"
echo "build_ref: wrote $HERE/_ref/libref.so"
