#!/bin/bash
# Builds oracle/libmgcm_oracle.so (C restatement, OpenMP) -- test infrastructure.
# -ffp-contract=off keeps every product/sum rounded separately, as the
# reference's x86-64 Fortran build does.
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
CC="${CC:-gcc}"
"$CC" -O2 -fPIC -shared -fopenmp -ffp-contract=off -fno-fast-math -std=c11 -Wall -Wno-maybe-uninitialized \
  -o "$HERE/libmgcm_oracle.so.tmp" "$HERE/mgcm_oracle.c" -lm
mv "$HERE/libmgcm_oracle.so.tmp" "$HERE/libmgcm_oracle.so"
echo "build_oracle: $HERE/libmgcm_oracle.so"
