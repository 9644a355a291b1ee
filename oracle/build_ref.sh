#!/usr/bin/env bash
# Builds the reference AMSS-NCKU Fortran of the BSSN RHS + RK4 path (and the
# Psi4 analysis routine getnp4) into
# oracle/_ref/libamssref.so (parity build: -O2 -ffp-contract=off) and
# oracle/_ref/libamssref_o3.so (timing build: -O3), straight from the
# reference sources where they lie.  Test infrastructure only.
#
# * Compiler: ROCm amdflang (the reference used gfortran, absent here).
# * Configuration: the reference's own checked-in macrodef.fh is what
#   `#include "macrodef.fh"` resolves to (same directory as the sources):
#   Cell, ghost_width 3, GAUGE 2, ABV 0.  Nothing is substituted for it.
# * fmisc.f90 is compiled from a filtered temporary copy with the one
#   subroutine `fwritefile` (fmisc.f90:1511-1526) removed: it opens a
#   DIRECT-access file without RECL (a gfortran extension) and amdflang
#   rejects it.  It is off the hot path; no other line is touched and the
#   temporary copy is deleted afterwards.  Nothing else is written.
set -euo pipefail
REF=${AMSS_REF_SRC:-/root/reference/amss-ncku-python/AMSS_NCKU_source}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
if [ ! -f "$REF/bssn_rhs.f90" ]; then
  echo "build_ref.sh: reference sources not found at $REF; skipping" >&2
  exit 0
fi
FC=${FC:-amdflang}
command -v "$FC" >/dev/null || { echo "build_ref.sh: $FC not found; skipping" >&2; exit 0; }
mkdir -p "$OUT"
TMP=$(mktemp -d)
trap 'rm -rf "$TMP"' EXIT
# fmisc.f90 minus fwritefile; the include of macrodef.fh is resolved with -I.
awk '/^ *subroutine fwritefile\(/{skip=1} !skip{print} /^ *end subroutine fwritefile/{skip=0}' \
    "$REF/fmisc.f90" > "$TMP/fmisc_nofwrite.f90"
# getnp4 (the Newman-Penrose Psi4) also holds the shell variant, which calls
# diff_new_sh's fderivs_shc: both are built so that the library loads whole
SRCS="bssn_rhs diff_new lopsidediff kodiss enforce_algebra rungekutta4_rout sommerfeld_rout prolongrestrict_cell getnp4 diff_new_sh"
build() {  # $1 = suffix, $2.. = flags
  local sfx=$1; shift
  local objs=()
  for s in $SRCS; do
    "$FC" -cpp "$@" -fPIC -I"$REF" -c "$REF/$s.f90" -o "$TMP/$s$sfx.o"
    objs+=("$TMP/$s$sfx.o")
  done
  "$FC" -cpp "$@" -fPIC -I"$REF" -c "$TMP/fmisc_nofwrite.f90" -o "$TMP/fmisc$sfx.o"
  objs+=("$TMP/fmisc$sfx.o")
  "$FC" -shared "$@" -o "$OUT/libamssref$sfx.so.tmp" "${objs[@]}"
  mv "$OUT/libamssref$sfx.so.tmp" "$OUT/libamssref$sfx.so"
}
build "" -O2 -ffp-contract=off
build "_o3" -O3
echo "build_ref.sh: built $OUT/libamssref.so and libamssref_o3.so"
