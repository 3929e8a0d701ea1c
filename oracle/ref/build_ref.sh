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
# mg_bintp_mom/tur are internal procedures of mg_driver (mg_driver.F90:332-333);
# a generated wrapper INCLUDEs them from the reference directory so the
# harness can call each one on its own.
cat > "$OUT/obj/harness_bintp.F90" <<EOF2
subroutine harness_bintp(which)
  use global
  use global_param
  use global_app
  use global_mpi
  use global_osg
  use global_turb
  implicit none
  integer, intent(in) :: which
  integer :: m, i, j, k, l
  if (which == 0) then
     call mg_bintp_mom
  else
     call mg_bintp_tur
  end if
contains
  include '$REF/mg_bintp_tur.F90'
  include '$REF/mg_bintp_mom.F90'
end subroutine harness_bintp
EOF2
cc "$OUT/obj/harness_bintp.F90" harness_bintp
# MPI_REAL arguments moved as 8-byte reals in this FP64 variant
gcc -O2 -c "$HERE/mpi_real8.c" -I"$MPI_INC" -o "$OUT/obj/mpi_real8.o"
cc "$HERE/desalm_ref_harness.F90" harness
REFOBJ=""
for o in precision global global_param global_app global_mpi global_osg global_turb checksum hooks_off \
         mg_driver solver_daf des_eddy mg_nlevel mg_inject mg_metrics bcond_fm rsmooth mg_brhs nonreflect_ibc \
         sgtsv xerbla harness_bintp mpi_real8 harness; do
    REFOBJ="$REFOBJ $OUT/obj/$o.o"
done
"$FC" -o "$OUT/desalm_ref_harness" $REFOBJ -L"$MPI_LIB" -lmpifort -lmpi -Wl,-rpath,"$MPI_LIB"
echo "built $OUT/desalm_ref_harness"

# The same harness with the reference's two case hooks linked in place of
# the empty bodies: conservation (src/conservation.F90, which INCLUDEs
# rhs_contra_j.F90 and rhs_exchng3_4d.F90 from its own directory) and the
# wall-function module (src/wf_mpi.F90).  Its module file goes to a
# directory of its own so the stub's wf_mpi.mod stays the one mg_driver saw.
mkdir -p "$OUT/obj_hooks"
"$FC" -O2 -fdefault-real-8 -ffp-contract=off -J"$OUT/obj_hooks" -I"$OUT/obj" -I"$MPI_INC" -c "$REF/wf_mpi.F90" -o "$OUT/obj_hooks/wf_mpi.o" \
    2>"$OUT/obj_hooks/wf_mpi.log" || { cat "$OUT/obj_hooks/wf_mpi.log" >&2; exit 1; }
cc "$REF/conservation.F90" conservation
"$FC" -o "$OUT/desalm_ref_harness_hooks" ${REFOBJ/ $OUT\/obj\/hooks_off.o/} \
    "$OUT/obj_hooks/wf_mpi.o" "$OUT/obj/conservation.o" -L"$MPI_LIB" -lmpifort -lmpi -Wl,-rpath,"$MPI_LIB"
echo "built $OUT/desalm_ref_harness_hooks"

# The same harness with the Fortran drop-in of des-alm_amd/fortran linked in
# place of one reference routine: desalm_shim_mgd swaps mg_driver,
# desalm_shim_daf swaps solver_daf (the reference mg_driver then calls the
# GPU solver_daf).  Needs des-alm_amd/libdesalm.so (found through $ORIGIN).
SHIM="$HERE/../../des-alm_amd/fortran"
LIBDIR="$HERE/../../des-alm_amd"
if [ -f "$LIBDIR/libdesalm.so" ]; then
    cc "$SHIM/desalm_f.F90" shim_desalm_f
    cc "$SHIM/mg_driver.F90" shim_mg_driver
    cc "$SHIM/solver_daf.F90" shim_solver_daf
    LIBS=(-L"$LIBDIR" -ldesalm -L"$MPI_LIB" -lmpifort -lmpi -Wl,-rpath,"$MPI_LIB" -Wl,-rpath,'$ORIGIN/../../des-alm_amd')
    "$FC" -o "$OUT/desalm_shim_mgd" ${REFOBJ/ $OUT\/obj\/mg_driver.o/} "$OUT/obj/shim_desalm_f.o" \
        "$OUT/obj/shim_mg_driver.o" "${LIBS[@]}"
    "$FC" -o "$OUT/desalm_shim_daf" ${REFOBJ/ $OUT\/obj\/solver_daf.o/} "$OUT/obj/shim_desalm_f.o" \
        "$OUT/obj/shim_solver_daf.o" "${LIBS[@]}"
    # solver_daf swapped inside the reference built with its own case hooks
    # (conservation, wf_mpi): the reference's mg_driver runs the hooks, the
    # drop-in the smoother passes
    HOOKOBJ="${REFOBJ/ $OUT\/obj\/hooks_off.o/}"
    "$FC" -o "$OUT/desalm_shim_daf_hooks" ${HOOKOBJ/ $OUT\/obj\/solver_daf.o/} "$OUT/obj_hooks/wf_mpi.o" \
        "$OUT/obj/conservation.o" "$OUT/obj/shim_desalm_f.o" "$OUT/obj/shim_solver_daf.o" "${LIBS[@]}"
    echo "built $OUT/desalm_shim_mgd $OUT/desalm_shim_daf $OUT/desalm_shim_daf_hooks"
fi
