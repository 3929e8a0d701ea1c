/*
 * mpi_real8.c -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
 *
 * The reference passes its reals to MPI as MPI_REAL (mg_residual.F90:71,
 * mg_bintp_mom.F90:134, mg_bintp_tur.F90:134, bcond_fm.F90:314,
 * conservation.F90:123,131).  Its FP64
 * variant, built by build_ref.sh with -fdefault-real-8, holds them in 8
 * bytes, so these Fortran MPI entry points pass MPI_REAL arguments on as
 * MPI_DOUBLE_PRECISION through MPICH's profiling interface (pmpi_*_).  The
 * calls of the one-rank harness that move data are exactly these four;
 * halo exchanges all go to MPI_PROC_NULL.
 */
#include <mpi.h>

extern void pmpi_alltoallv_(void*, MPI_Fint*, MPI_Fint*, MPI_Fint*, void*, MPI_Fint*, MPI_Fint*, MPI_Fint*,
                            MPI_Fint*, MPI_Fint*);
extern void pmpi_reduce_(void*, void*, MPI_Fint*, MPI_Fint*, MPI_Fint*, MPI_Fint*, MPI_Fint*, MPI_Fint*);
extern void pmpi_bcast_(void*, MPI_Fint*, MPI_Fint*, MPI_Fint*, MPI_Fint*, MPI_Fint*);
extern void pmpi_allreduce_(void*, void*, MPI_Fint*, MPI_Fint*, MPI_Fint*, MPI_Fint*, MPI_Fint*);

static MPI_Fint real8(const MPI_Fint* t) {
    return *t == MPI_Type_c2f(MPI_REAL) ? MPI_Type_c2f(MPI_DOUBLE_PRECISION) : *t;
}

void mpi_alltoallv_(void* sb, MPI_Fint* sc, MPI_Fint* sd, MPI_Fint* st, void* rb, MPI_Fint* rc, MPI_Fint* rd,
                    MPI_Fint* rt, MPI_Fint* comm, MPI_Fint* ierr) {
    MPI_Fint s = real8(st), r = real8(rt);
    pmpi_alltoallv_(sb, sc, sd, &s, rb, rc, rd, &r, comm, ierr);
}

void mpi_reduce_(void* sb, void* rb, MPI_Fint* count, MPI_Fint* t, MPI_Fint* op, MPI_Fint* root, MPI_Fint* comm,
                 MPI_Fint* ierr) {
    MPI_Fint tt = real8(t);
    pmpi_reduce_(sb, rb, count, &tt, op, root, comm, ierr);
}

void mpi_bcast_(void* buf, MPI_Fint* count, MPI_Fint* t, MPI_Fint* root, MPI_Fint* comm, MPI_Fint* ierr) {
    MPI_Fint tt = real8(t);
    pmpi_bcast_(buf, count, &tt, root, comm, ierr);
}

void mpi_allreduce_(void* sb, void* rb, MPI_Fint* count, MPI_Fint* t, MPI_Fint* op, MPI_Fint* comm, MPI_Fint* ierr) {
    MPI_Fint tt = real8(t);
    pmpi_allreduce_(sb, rb, count, &tt, op, comm, ierr);
}
