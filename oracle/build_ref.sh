#!/bin/bash
# Test infrastructure only: builds the reference Fortran solver (FP64) and the
# dump driver ref_driver.f90 into oracle/_ref/.  Nothing here is shipped or
# measured as the product.
#
# The reference does not compile as shipped (names missing from USE ... ONLY
# lists, an array passed to a scalar dummy, a DATA statement on a
# use-associated IAXIS).  The patch below is applied to a throw-away copy in a
# temporary directory that is deleted on exit; it changes no arithmetic:
#   * geometry.f90   : import N from SOLVER_STATE
#   * viscous.f90    : import I, CP (SOLVER_STATE) and HALF (CONSTANTS)
#   * boundary.f90   : import HALF; pass NCk(NP) (a scalar) to
#                      ACCUMULATE_GPORT_ALONG_BOUNDARY
#   * residual.f90   : import RTOLERANCE and HALF
#   * source_terms.f90: import FR2
#   * all five physics modules: drop the private, never-assigned
#     "INTEGER :: IAXIS" and use SOLVER_STATE's IAXIS (the copy flow.in sets)
#   * turbnstwod.f90 : drop "DATA IAXIS" (flow.in overrides it anyway)
#   * NN = 60000 (the shipped value) for turbnstwod / ref_driver, and a second
#     build with NN = REF_NN_BIG (default 1060000, enough for the 2049x513
#     ramp's cells + ghosts; 0 skips it) for turbnstwod_big / ref_driver_big
# Build flags: -O3 only.  The reference Makefile's -fdefault-real-8 promotes
# DOUBLE PRECISION to REAL(16); the oracle is the FP64 build.
set -euo pipefail
REF=${TDT_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT=$HERE/_ref
FC=${FC:-/opt/rocm/bin/amdflang}
if [ ! -d "$REF/src/solvers" ]; then
    echo "build_ref.sh: $REF/src/solvers not present; reference oracle not built"
    exit 0
fi
mkdir -p "$OUT"
TMP=$(mktemp -d /tmp/tdt_refbuild.XXXXXX)
trap 'rm -rf "$TMP"' EXIT

build_variant() {   # $1 = NN value, $2 = suffix
    local nn=$1 sfx=$2
    local d=$TMP/b$sfx
    mkdir -p "$d"
    for f in constants solver_state geometry viscous source_terms boundary residual turbnstwod; do
        cp "$REF/src/solvers/$f.f90" "$d/$f.f90"
    done
    sed -i "s/INTEGER, PARAMETER :: NN = 60000/INTEGER, PARAMETER :: NN = $nn/" "$d/constants.f90"
    sed -i 's/YCCFACE1, YCCFACE2, YCCFACE3, YCCFACE4, I$/YCCFACE1, YCCFACE2, YCCFACE3, YCCFACE4, I, N, IAXIS/' "$d/geometry.f90"
    sed -i 's/RTRED, YAVG, VBYY$/RTRED, YAVG, VBYY, I, CP, IAXIS/' "$d/viscous.f90"
    sed -i 's/QUARTER, YAXIS_THRESH$/QUARTER, YAXIS_THRESH, HALF/' "$d/viscous.f90"
    sed -i 's/AMUINIT, RINIT, N, I$/AMUINIT, RINIT, N, I, IAXIS/' "$d/boundary.f90"
    sed -i 's/USE CONSTANTS, ONLY: PA_TO_BAR, BL_EXPONENT, GTH_VAL$/USE CONSTANTS, ONLY: PA_TO_BAR, BL_EXPONENT, GTH_VAL, HALF/' "$d/boundary.f90"
    sed -i -E 's/CALL ACCUMULATE_GPORT_ALONG_BOUNDARY\((NC[1-4]), ([1-4])\)/CALL ACCUMULATE_GPORT_ALONG_BOUNDARY(\1(NP), \2)/' "$d/boundary.f90"
    sed -i 's/G4I, G5I, DTBYVOL, ALPHA1, CFLTURB, I$/G4I, G5I, DTBYVOL, ALPHA1, CFLTURB, I, RTOLERANCE, IAXIS/' "$d/residual.f90"
    sed -i 's/USE CONSTANTS, ONLY: CFLTURB_VAL$/USE CONSTANTS, ONLY: CFLTURB_VAL, HALF/' "$d/residual.f90"
    sed -i 's/CONST2, CONST3, CMU, VBYY$/CONST2, CONST3, CMU, VBYY, FR2, IAXIS/' "$d/source_terms.f90"
    for f in geometry viscous source_terms boundary residual; do
        sed -i '/^    INTEGER :: IAXIS  ! 0 = 2D Cartesian, 1 = Axisymmetric$/d' "$d/$f.f90"
    done
    sed -i '/^    DATA IAXIS \/ 0 \/$/d' "$d/turbnstwod.f90"
    # COMPUTE_PRIMITIVE_VARS is an external subroutine in the main-program
    # file; the dump driver links it from its own object.
    sed -n '/^SUBROUTINE COMPUTE_PRIMITIVE_VARS/,/^END SUBROUTINE COMPUTE_PRIMITIVE_VARS/p' \
        "$d/turbnstwod.f90" > "$d/primvars.f90"
    ( cd "$d"
      for f in constants solver_state geometry viscous source_terms boundary residual primvars; do
          "$FC" -O3 -c "$f.f90"
      done
      "$FC" -O3 -c turbnstwod.f90
      "$FC" -O3 -o "$OUT/turbnstwod$sfx" constants.o solver_state.o geometry.o viscous.o \
          source_terms.o boundary.o residual.o turbnstwod.o
      "$FC" -O3 -c "$HERE/ref_driver.f90"
      "$FC" -O3 -o "$OUT/ref_driver$sfx" constants.o solver_state.o geometry.o viscous.o \
          source_terms.o boundary.o residual.o primvars.o ref_driver.o )
}

build_variant 60000 ""
REF_NN_BIG=${REF_NN_BIG:-1060000}
if [ "$REF_NN_BIG" != "0" ]; then
    build_variant "$REF_NN_BIG" "_big"
fi
echo "build_ref.sh: built $(ls "$OUT" | tr '\n' ' ')"
