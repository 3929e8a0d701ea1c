#!/bin/bash
# Builds oracle/_ref/desalm_ref_harness: the reference's own Fortran sources
# for the inner iteration, compiled where they lie under
# /root/reference/src (never copied), in FP64 (precision_fp64.F90 +
# -fdefault-real-8), with the wall-function and outlet-conservation hooks
# switched off (hooks_off.F90), linked against the image's MPICH.
# Test infrastructure only; outputs go to oracle/_ref/ and nowhere else.
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
REF="${DESALM_REFERENCE_SRC:-/root/reference/src}"
OUT="$HERE/../_ref"
FC="${FC:-/opt/rocm/bin/amdflang}"
MPI_INC="${MPI_INC:-/opt/conda/include}"
MPI_LIB="${MPI_LIB:-/opt/conda/lib}"
if [ ! -d "$REF" ]; then
    echo "reference sources not found at $REF; skipping oracle/_ref build" >&2
    exit 0
fi
mkdir -p "$OUT/obj"
# No -I for the reference directory: it also holds shipped compiler output
# (.mod/.o) that must never be read; Fortran INCLUDE lines resolve against
# the including file's own directory.
FFLAGS="-O2 -fdefault-real-8 -ffp-contract=off -J$OUT/obj -I$MPI_INC"
cc() { "$FC" $FFLAGS -c "$1" -o "$OUT/obj/$2.o" 2>"$OUT/obj/$2.log" || { cat "$OUT/obj/$2.log" >&2; exit 1; }; }
cc "$HERE/precision_fp64.F90" precision
for m in global global_param global_app global_mpi global_osg global_turb checksum; do
    cc "$REF/$m.F90" "$m"
done
cc "$HERE/hooks_off.F90" hooks_off
for u in mg_driver solver_daf des_eddy mg_nlevel mg_inject mg_metrics bcond_fm rsmooth mg_brhs nonreflect_ibc; do
    cc "$REF/$u.F90" "$u"
done
cc "$REF/sgtsv.f" sgtsv
cc "$REF/xerbla.f" xerbla
cc "$HERE/desalm_ref_harness.F90" harness
"$FC" -o "$OUT/desalm_ref_harness" "$OUT"/obj/*.o -L"$MPI_LIB" -lmpifort -lmpi -Wl,-rpath,"$MPI_LIB"
echo "built $OUT/desalm_ref_harness"
