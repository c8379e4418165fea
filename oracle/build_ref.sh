#!/bin/bash
# Compiles the reference's FP64 leaf kernels, from the PICSAR sources where
# they lie, into oracle/_ref/libpxrref.so (test infrastructure only).
#
# Nothing from the reference tree is copied into the repository: routines
# are cut out of the reference files into a temporary directory, compiled
# there, and only the shared library is kept.  Routines are extracted by
# name because three of the source files cannot be compiled whole outside
# the reference's build (they USE mpi and module state, or carry an
# unrelated routine that flang rejects):
#   modules.F90           -> modules PICSAR_precision, constants, precomputed
#   direct_current_deposition_3d.F90 -> depose_jxjyjz_scalar_{1_1_1,2_2_2,3_3_3}
#   sorting.F90           -> pxr_particle_bin_sorting
#   submain.F90           -> FD_weights
#   laser_pusher_manager_3d.F90 -> gaussian_profile, hanning_profile (the
#                            pushers around them read module state)
#   diags.F90             -> calc_field_div, calc_field_divB, placed in a
#                            module that gives them the host context of
#                            their own module (USE PICSAR_precision,
#                            constants; the MPI module it also uses is not
#                            needed by these two)
# -DLIBRARY selects normalised units (clight = mu0 = eps0 = 1).
#
# Two libraries are built from the same sources:
#   libpxrref.so     -O2 -ffp-contract=off: the golden vectors (no contraction,
#                    so that the oracle's operation order can match it);
#   libpxrref_o3.so  -O3 -fopenmp, the reference's production flags
#                    (Makefile:56, without -g -fbounds-check): the CPU baseline.
#                    The leaf kernels hold !$OMP SIMD directives only, which
#                    -fopenmp turns on; threads run over tiles, as the
#                    reference's managers do.
set -euo pipefail
REF=${PXR_REF:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
FC=${FC:-amdflang}
FFLAGS_GOLDEN="-cpp -DLIBRARY -O2 -fPIC -ffp-contract=off"
FFLAGS_PROD="-cpp -DLIBRARY -O3 -fopenmp -fPIC"
SRC="$REF/src"
if [ ! -d "$SRC" ]; then
  echo "build_ref: reference tree not found at $REF" >&2
  exit 2
fi
command -v "$FC" >/dev/null || { echo "build_ref: $FC not found" >&2; exit 2; }
TMP=$(mktemp -d)
trap 'rm -rf "$TMP"' EXIT

# extract "MODULE first ... END MODULE last" (inclusive)
awk 'BEGIN{IGNORECASE=1} /^ *MODULE +PICSAR_precision/{on=1} on{print} /^ *END MODULE +precomputed/{if(on) exit}' \
  "$SRC/modules/modules.F90" > "$TMP/pxr_consts.F90"

# extract one SUBROUTINE by name (up to its END SUBROUTINE)
extract_sub() { # file name
  awk -v name="$2" 'BEGIN{IGNORECASE=1; pat="^ *SUBROUTINE +" name " *\\("}
    $0 ~ pat {on=1}
    on {print}
    on && /^ *END +SUBROUTINE/ {exit}' "$1"
}
D="$SRC/particle_deposition/current_deposition/direct/direct_current_deposition_3d.F90"
: > "$TMP/direct_scalar.F90"
for o in 1_1_1 2_2_2 3_3_3; do extract_sub "$D" "depose_jxjyjz_scalar_$o" >> "$TMP/direct_scalar.F90"; done
extract_sub "$SRC/housekeeping/sorting.F90" pxr_particle_bin_sorting > "$TMP/binsort.F90"
extract_sub "$SRC/submain.F90" FD_weights > "$TMP/fdweights.F90"
L="$SRC/particle_pushers/laser_pusher_manager_3d.F90"
{ extract_sub "$L" gaussian_profile; extract_sub "$L" hanning_profile; } > "$TMP/laserprof.F90"
G="$SRC/diags/diags.F90"
{ printf 'MODULE pxr_ref_diags\n  USE PICSAR_precision\n  USE constants\n  IMPLICIT NONE\nCONTAINS\n'
  extract_sub "$G" calc_field_div; extract_sub "$G" calc_field_divB
  printf 'END MODULE pxr_ref_diags\n'; } > "$TMP/divdiags.F90"

cd "$TMP"
# compile every unit with the flags $1 into the shared library $2
build_lib() {
  local FFLAGS="$1" LIB="$2" LDF="$3"
  mkdir -p "$TMP/$LIB.d"
  cd "$TMP/$LIB.d"
  $FC $FFLAGS -c ../pxr_consts.F90 -o pxr_consts.o
  local OBJS="pxr_consts.o"
  for f in field_gathering/energy_conserving/field_gathering_o1_3d.F90 \
           field_gathering/energy_conserving/field_gathering_on_3d.F90 \
           field_gathering/energy_conserving/field_gathering_o2_3d.F90 \
           field_gathering/energy_conserving/field_gathering_o3_3d.F90 \
           particle_pushers/boris_pusher/boris_3d.F90 \
           particle_pushers/vay_pusher/vay_3d.F90 \
           particle_deposition/current_deposition/esirkepov/esirkepov_3d.F90 \
           particle_deposition/charge_deposition/charge_deposition_3d.F90 \
           field_solvers/Maxwell/yee_solver/yee.F90; do
    local o=$(basename "$f" .F90).o
    $FC $FFLAGS -c "$SRC/$f" -o "$o"
    OBJS="$OBJS $o"
  done
  for f in direct_scalar binsort fdweights laserprof divdiags; do
    $FC $FFLAGS -c "../$f.F90" -o "$f.o"
    OBJS="$OBJS $f.o"
  done
  mkdir -p "$OUT"
  $FC -shared $LDF -o "$OUT/$LIB.tmp" $OBJS
  mv "$OUT/$LIB.tmp" "$OUT/$LIB"
  echo "build_ref: wrote $OUT/$LIB"
  cd "$TMP"
}
build_lib "$FFLAGS_GOLDEN" libpxrref.so "" &
build_lib "$FFLAGS_PROD" libpxrref_o3.so "-fopenmp" &
wait %1
wait %2
# the CPU baseline's OpenMP tile loop over the production kernels
CC=${CC_OMP:-/opt/rocm/llvm/bin/clang}
$CC -O2 -fopenmp -fPIC -shared "$HERE/ref_tile_driver.c" -L"$OUT" -lpxrref_o3 \
    -Wl,-rpath,'$ORIGIN' -o "$OUT/libpxrref_driver.so.tmp"
mv "$OUT/libpxrref_driver.so.tmp" "$OUT/libpxrref_driver.so"
echo "build_ref: wrote $OUT/libpxrref_driver.so"
