#!/bin/bash
# Builds oracle/_ref/libmgcm_ref.so: the reference Fortran radiation routines,
# compiled from their own sources where they lie under /root/reference/code,
# linked with oracle/ref_harness.f90 (bind(C) driver), and oracle/_ref/htest,
# the reference's history-file inspection tool.  Test infrastructure only.
# Objects and module files go to a temporary directory; only the shared library
# and htest land in oracle/_ref/.  Exits 0 without building when the reference
# tree is absent.
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
REF="${MGCM_REFERENCE_DIR:-/root/reference}/code"
OUT="$HERE/_ref"
if [ ! -d "$REF" ]; then
  echo "build_ref: reference sources not present; skipping" >&2
  exit 0
fi
FC="${FC:-/opt/rocm/bin/amdflang}"
TMP="$(mktemp -d "${TMPDIR:-/tmp}/mgcm_ref.XXXXXX")"
trap 'rm -rf "$TMP"' EXIT
mkdir -p "$OUT"
SRCS="modules.f90 tpindex.f90 optcv.f90 optci.f90 getdetau.f90 gfluxv.f90 gfluxi.f90
      dsolver.f90 dtridgl.f90 sfluxv.f90 sfluxi.f90 setspv.f90 setspi.f90 setrad.f90
      laginterp.f90 lagrange.f90 initcld.f opt_dst.f opt_cld.f dustprofile.f filltaucum.f
      jsrchgt.f90 dsolflux.f fillpt.f90"
cd "$TMP"
# -O2 without fast-math; x86-64 baseline (no FMA), as the reference's own gfortran -O3 build.
for f in $SRCS; do
  "$FC" -O2 -fPIC -w -c "$REF/$f" -o "${f%.*}.o"
done
"$FC" -O2 -fPIC -c "$HERE/ref_harness.f90" -o ref_harness.o
"$FC" -shared -fPIC -o "$OUT/libmgcm_ref.so.tmp" ./*.o
mv "$OUT/libmgcm_ref.so.tmp" "$OUT/libmgcm_ref.so"
echo "build_ref: $OUT/libmgcm_ref.so"
# The reference's history-file reader and its point-inspection tool, as shipped
# (tutorial/historymod.f90, tutorial/htest.f90): the check that history files
# written by the library open in the reference's analysis tools.
TUT="$REF/../tutorial"
mkdir -p "$TMP/htest"
cd "$TMP/htest"
"$FC" -O2 -w -c "$TUT/historymod.f90" -o historymod.o
"$FC" -O2 -w -c "$TUT/htest.f90" -o htest.o
"$FC" -o "$OUT/htest.tmp" htest.o historymod.o
mv "$OUT/htest.tmp" "$OUT/htest"
echo "build_ref: $OUT/htest"
