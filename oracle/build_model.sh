#!/bin/bash
# Builds the whole reference GCM (test infrastructure for the in-model drop-in
# test), three ways, into oracle/_ref/:
#   gcm_ref      the reference's own sources as they are;
#   gcm_refglue  COMP3 with its four radiation calls moved into the glue routines
#                mgcm_comp3_vis / mgcm_comp3_ir of oracle/comp3_ref_glue.f90,
#                which run the reference's OPTCV+SFLUXV and OPTCI+SFLUXI (the
#                same calls, the same arguments) -- behaves as gcm_ref, bit for bit;
#   gcm_mi355x   the same rewritten COMP3 with the library's glue
#                (legacy-mars-global-climate-model_amd/fortran/comp3_rt.f90): the
#                radiation on the MI355X through mgcm_rt_mod / libmgcm_rt.so,
#                one column per call;
#   gcm_refsplit COMP3's column loop in two passes around one radiation batch
#                (fortran/comp3_split.f90), the batch run by the reference's own
#                routines column by column (oracle/comp3_ref_split.f90) --
#                behaves as gcm_ref, bit for bit;
#   gcm_mi355x_split  the same two-pass COMP3, the batch one mgcm_rt_radiation
#                call on the MI355X (fortran/comp3_rt_batch.f90).
# Every glue build writes every radiation call's outputs to $MGCM_COMP3_DUMP when
# it is set.  Also copies run/data and writes a one-hour tutorial namelist
# (oracle/_ref/gcm_run/), which the test runs in its own directories.
#
# Every build compiles the reference's code/ sources with amdflang -O2, as they
# lie under the reference tree, with two edits made in a temporary copy only:
#   main.f   the ENDFILE of unit 6 after the last time step is dropped (flang's
#            runtime stops the program on the write that follows it; gfortran
#            accepts it);
#   comp3.f  (glue builds) the calls OPTCV (code/comp3.f:658-662) and SFLUXV
#            (:672-675) become one call of mgcm_comp3_vis, OPTCI (:724-727) and
#            SFLUXI (:731-733) one call of mgcm_comp3_ir; the lines are checked
#            before they are replaced.  (split builds) the column loop (DO 500 /
#            DO 490, :411-1044) becomes two loops: the first runs :417-636, the
#            albedos (:642-654 under the sun-up test, :712-720), records the
#            column's radiation inputs and saves its state; after the batch the
#            second restores each column and runs :637-1042 with the glue calls;
#            TEMPGR's DSOLFLUX loop (:311-351) records each column's ACOSZ and
#            one call (mgcm_comp3_tempgr) forms every DNVFLUX.
# The rewritten sources live in a temporary directory and are deleted; only the
# linked programs land in oracle/_ref/.  Exits 0 without building when the
# reference tree is absent.
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
ROOT="$(dirname "$HERE")"
REF="${MGCM_REFERENCE_DIR:-/root/reference}"
CODE="$REF/code"
OUT="$HERE/_ref"
PKG="$ROOT/legacy-mars-global-climate-model_amd"
LIBDIR="$PKG/mgcm_rt"
if [ ! -d "$CODE" ]; then
  echo "build_model: reference sources not present; skipping" >&2
  exit 0
fi
if [ ! -f "$LIBDIR/libmgcm_rt.so" ]; then
  echo "build_model: $LIBDIR/libmgcm_rt.so not built (run legacy-mars-global-climate-model_amd/build.py)" >&2
  exit 1
fi
FC="${FC:-/opt/rocm/bin/amdflang}"
JOBS="${MAX_JOBS:-8}"
TMP="$(mktemp -d "${TMPDIR:-/tmp}/mgcm_model.XXXXXX")"
trap 'rm -rf "$TMP"' EXIT
mkdir -p "$OUT" "$TMP/obj"

# the model's source list, from its Makefile
SRCS=$(sed -n '/^SRC *=/,/modules\.f90 *$/p' "$CODE/Makefile" | sed 's/^SRC *=//; s/\\//g')

# main.f without the final ENDFILE of unit 6
grep -q -E '^ +ENDFILE MTP *$' "$CODE/main.f"
grep -v -E '^ +ENDFILE MTP *$' "$CODE/main.f" > "$TMP/main.f"

# comp3.f with the radiation calls replaced by the glue calls
check() {  # line number, expected call
  sed -n "${1}p" "$CODE/comp3.f" | grep -q -i "call $2(" || { echo "build_model: comp3.f:$1 is not 'call $2'" >&2; exit 1; }
}
check 658 optcv
check 672 sfluxv
check 724 optci
check 731 sfluxi
awk '
NR == 658 {
  print "                 call mgcm_comp3_vis(PLEV,TLEV,PMID,TMID,QH2O,TAUREF,"
  print "     *                TAUREFCLD,ALS,ACOSZ,SOL,TAUCUMV,NFLUXTOPV,FMNETV,"
  print "     *                fluxupv,fluxdnv,diffvt,jcmn,icmn)"
}
NR == 724 {
  print "               call mgcm_comp3_ir(PLEV,TLEV,PMID,TMID,QH2O,TAUREF,"
  print "     *                TAUREFCLD,ALBI,NFLUXTOPI,FMNETI,FLUXUPI,FLUXDNI,"
  print "     *                jcmn,icmn)"
}
(NR >= 658 && NR <= 662) || (NR >= 672 && NR <= 675) || (NR >= 724 && NR <= 727) || (NR >= 731 && NR <= 733) { next }
{ print }' "$CODE/comp3.f" > "$TMP/comp3_glue.f"
# the same rewrite keeping the original line numbers (each replaced line blank,
# the glue call on one long line), for the split build's line ranges
awk '
NR == 658 { print "                 call mgcm_comp3_vis(PLEV,TLEV,PMID,TMID,QH2O,TAUREF,"; next }
NR == 659 { print "     *                TAUREFCLD,ALS,ACOSZ,SOL,TAUCUMV,NFLUXTOPV,FMNETV,"; next }
NR == 660 { print "     *                fluxupv,fluxdnv,diffvt,jcmn,icmn)"; next }
NR == 724 { print "               call mgcm_comp3_ir(PLEV,TLEV,PMID,TMID,QH2O,TAUREF,"; next }
NR == 725 { print "     *                TAUREFCLD,ALBI,NFLUXTOPI,FMNETI,FLUXUPI,FLUXDNI,"; next }
NR == 726 { print "     *                jcmn,icmn)"; next }
(NR >= 658 && NR <= 662) || (NR >= 672 && NR <= 675) || (NR >= 724 && NR <= 727) || (NR >= 731 && NR <= 733) { print ""; next }
{ print }' "$CODE/comp3.f" > "$TMP/comp3_glue_full.f"

# comp3.f in two passes around one radiation batch
grep -q -E '^ +DO 500 I=1,IM *$' <(sed -n 411p "$CODE/comp3.f") || { echo "build_model: comp3.f:411 is not DO 500" >&2; exit 1; }
grep -q -E '^ +DO 490 J=1,JM-1 *$' <(sed -n 415p "$CODE/comp3.f") || { echo "build_model: comp3.f:415 is not DO 490" >&2; exit 1; }
grep -q -E '^ +if\(acosz.ge.1.0e-5\) then *$' <(sed -n 637p "$CODE/comp3.f") || { echo "build_model: comp3.f:637" >&2; exit 1; }
grep -q -E '^ +ALS += ALSP' <(sed -n 642p "$CODE/comp3.f") || { echo "build_model: comp3.f:642" >&2; exit 1; }
grep -q -E '^ +surfalb\(J,I\) = ALS *$' <(sed -n 654p "$CODE/comp3.f") || { echo "build_model: comp3.f:654" >&2; exit 1; }
grep -q -E '^ +ALBI += 1.0D0-EGOGND *$' <(sed -n 712p "$CODE/comp3.f") || { echo "build_model: comp3.f:712" >&2; exit 1; }
grep -q -E '^ +ENDIF *$' <(sed -n 720p "$CODE/comp3.f") || { echo "build_model: comp3.f:720" >&2; exit 1; }
grep -q -E '^  490 +CONTINUE *$' <(sed -n 1043p "$CODE/comp3.f") || { echo "build_model: comp3.f:1043" >&2; exit 1; }
grep -q -E '^  500 +CONTINUE *$' <(sed -n 1044p "$CODE/comp3.f") || { echo "build_model: comp3.f:1044" >&2; exit 1; }
grep -q -E '^ +if\(acosz.ge.1.0e-5\) then *$' <(sed -n 335p "$CODE/comp3.f") || { echo "build_model: comp3.f:335" >&2; exit 1; }
grep -q -E '^ +DNVFLUX\(J,I\) += DNDIFFV\(J,I\) \+ DIRECTSOL *$' <(sed -n 348p "$CODE/comp3.f") || { echo "build_model: comp3.f:348" >&2; exit 1; }
grep -q -E '^ +END DO +! I-loop' <(sed -n 351p "$CODE/comp3.f") || { echo "build_model: comp3.f:351" >&2; exit 1; }
{
  sed -n '1,334p' "$CODE/comp3.f"
  echo "          call mgcm_comp3_tempgr_record(ACOSZ,J,I)"
  sed -n '349,351p' "$CODE/comp3.f"
  echo "      call mgcm_comp3_tempgr(SOL,DNVFLUX)"
  sed -n '352,410p' "$CODE/comp3.f"
  echo "         call comp3_split_pass1_begin()"
  sed -n '411,636p' "$CODE/comp3.f"
  echo "               if(acosz.ge.1.0e-5) then"
  sed -n '642,654p' "$CODE/comp3.f"
  echo "               end if"
  sed -n '712,720p' "$CODE/comp3.f"
  echo "               call mgcm_comp3_record(PLEV,TLEV,PMID,TMID,QH2O,"
  echo "     *              TAUREF,TAUREFCLD,ALS,ALBI,ACOSZ,SOL,jcmn,icmn)"
  echo "               call comp3_split_save(jcmn,icmn,QPI,QPIG,TETAORG,"
  echo "     *              QPISAV,dustref,PLEV,TLEV,PMID,TMID,QH2O)"
  sed -n '1043,1044p' "$CODE/comp3.f"
  echo "         call comp3_split_pass1_end()"
  echo "         call comp3_split_radiate()"
  echo "         DO 501 I=1,IM"
  echo "            DO 491 J=1,JM-1"
  echo "               call comp3_split_restore(J,I,QPI,QPIG,TETAORG,"
  echo "     *              QPISAV,dustref,PLEV,TLEV,PMID,TMID,QH2O)"
  awk 'NR >= 637 && NR <= 1042' "$TMP/comp3_glue_full.f"
  echo "  491       CONTINUE"
  echo "  501    CONTINUE"
  sed -n '1045,$p' "$CODE/comp3.f"
} > "$TMP/comp3_split.f"

cd "$TMP/obj"
FFLAGS="-O2 -w -J $TMP/obj"
"$FC" $FFLAGS -c "$CODE/modules.f90" -o modules.o
# every other source as it is (main.f and comp3.f from their copies)
printf '%s\n' $SRCS | grep -v -E '^(modules\.f90|main\.f|comp3\.f)$' |
  xargs -P "$JOBS" -I{} sh -c '"$0" $1 -c "$2/{}" -o "$(basename {} | sed "s/\.[^.]*$//").o"' "$FC" "$FFLAGS" "$CODE"
"$FC" $FFLAGS -c "$TMP/main.f" -o main.o
"$FC" $FFLAGS -c "$CODE/comp3.f" -o "$TMP/comp3_orig.o"
"$FC" $FFLAGS -c "$TMP/comp3_glue.f" -o "$TMP/comp3_glue.o"
"$FC" $FFLAGS -c "$HERE/comp3_ref_glue.f90" -o "$TMP/comp3_ref_glue.o"
"$FC" $FFLAGS -c "$PKG/fortran/mgcm_rt_mod.f90" -o "$TMP/mgcm_rt_mod.o"
"$FC" $FFLAGS -c "$PKG/fortran/comp3_rt_ctx.f90" -o "$TMP/comp3_rt_ctx.o"
"$FC" $FFLAGS -c "$PKG/fortran/comp3_rt.f90" -o "$TMP/comp3_rt.o"
"$FC" $FFLAGS -c "$PKG/fortran/comp3_split.f90" -o "$TMP/comp3_split.o"
"$FC" $FFLAGS -c "$PKG/fortran/comp3_rt_batch.f90" -o "$TMP/comp3_rt_batch.o"
"$FC" $FFLAGS -c "$HERE/comp3_ref_split.f90" -o "$TMP/comp3_ref_split.o"
"$FC" $FFLAGS -c "$TMP/comp3_split.f" -o "$TMP/comp3_splitloop.o"
COMMON=$(ls *.o)
RPATH='$ORIGIN/../../legacy-mars-global-climate-model_amd/mgcm_rt'
"$FC" -o "$OUT/gcm_ref.tmp" $COMMON "$TMP/comp3_orig.o"
"$FC" -o "$OUT/gcm_refglue.tmp" $COMMON "$TMP/comp3_glue.o" "$TMP/comp3_ref_glue.o"
"$FC" -o "$OUT/gcm_mi355x.tmp" $COMMON "$TMP/comp3_glue.o" "$TMP/comp3_rt.o" "$TMP/comp3_rt_ctx.o" \
  "$TMP/mgcm_rt_mod.o" -L "$LIBDIR" -lmgcm_rt -Wl,-rpath,"$RPATH"
"$FC" -o "$OUT/gcm_refsplit.tmp" $COMMON "$TMP/comp3_splitloop.o" "$TMP/comp3_split.o" "$TMP/comp3_ref_split.o"
"$FC" -o "$OUT/gcm_mi355x_split.tmp" $COMMON "$TMP/comp3_splitloop.o" "$TMP/comp3_split.o" \
  "$TMP/comp3_rt_batch.o" "$TMP/comp3_rt_ctx.o" "$TMP/mgcm_rt_mod.o" -L "$LIBDIR" -lmgcm_rt -Wl,-rpath,"$RPATH"
for b in gcm_ref gcm_refglue gcm_mi355x gcm_refsplit gcm_mi355x_split; do mv "$OUT/$b.tmp" "$OUT/$b"; done

# run directory: the model's data files and the tutorial namelist cut to one hour
# of model time (30 steps of 2 minutes, COMP3's radiation every 8 steps),
# history every half hour
RUN="$OUT/gcm_run"
mkdir -p "$RUN/data"
cp "$REF"/run/data/* "$RUN/data/"
grep -q -E '^ +taue *= *[0-9.]+ +tauh *= *[0-9.]+' "$REF/tutorial/mars_tutorial"
sed -E 's/^( +)taue *= *[0-9.]+ +tauh *= *[0-9.]+/\1taue = 1.0  tauh = 0.5/' "$REF/tutorial/mars_tutorial" \
  > "$RUN/mars_1h"
echo "build_model: gcm_ref gcm_refglue gcm_mi355x gcm_refsplit gcm_mi355x_split in $OUT, run directory $RUN"
