/*
 * desalm_oracle.h -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
 *
 * CPU FP64 restatement of the reference's pseudo-time inner iteration, used
 * only as the checker by tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg.  Nothing in des-alm_amd/ may include, link or call it.
 *
 * The oracle keeps the reference's own module-state layout: arrays over all
 * multigrid levels concatenated (le_idx numbering, src/init.F90:880-980),
 * array-of-structures components (q(1:5,l), rh(1:4,l), csi(1:3,l)).
 */
#ifndef DESALM_ORACLE_H
#define DESALM_ORACLE_H

#include "../include/desalm.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_state or_state;

or_state* or_create(const desalm_params* p);
void or_destroy(or_state* o);
long or_nall(const or_state* o);
long or_level_offset(const or_state* o, int n);

/* mg_inject(x,y,z) + mg_metrics on all levels (src/mg_metrics.F90:56-300). */
int or_set_grid(or_state* o, const double* x, const double* y, const double* z);
/* fine-level wall distance; wd of the coarse levels stays zero (init.F90:2313) */
int or_set_wd(or_state* o, const double* wd_fine);
/* fine-level state (AoS q(1:5,n1) etc.); NULL leaves a field as it is */
int or_set_state(or_state* o, const double* q, const double* qn, const double* qnm1,
                 const double* xnut);
int or_get_state(const or_state* o, double* q, double* qn, double* qnm1, double* xnut);

/* whole-array access, all levels, reference layout.
 * names: q qn qnm1 qold_mg (5) rh pk (4) csi eta zet (3) aj xnut wd dtev x y z (1) */
int or_get_array(const or_state* o, const char* name, double* out);
int or_set_array(or_state* o, const char* name, const double* in);

/* one physical step of mg_driver (src/mg_driver.F90:2-323);
 * resid: itm*5 log10 norms (may be NULL) */
int or_mg_driver(or_state* o, int* itc, double* resid);
/* one DAF pass on level n of the stored state (src/solver_daf.F90) */
int or_solver_daf(or_state* o, int n, int decide_recalc_rh, int decide_calc_pk);
/* residual assembly only on level n (rh of solver_daf before the DAF) */
int or_residual(or_state* o, int n, int calc_pk);
int or_dtau(or_state* o, int n, double* out);
/* one des_eddy call on the fine level of the stored state (src/des_eddy.F90) */
int or_des_eddy(or_state* o);
/* mg_nlevel(n) + mg_zero_pk_boundary (src/mg_nlevel.F90:2-150) */
int or_mg_nlevel(or_state* o, int n);
/* mg_prolong(n) (src/mg_prolong.F90:1-281) */
int or_mg_prolong(or_state* o, int n);
/* mg_residual(1) (src/mg_residual.F90:2-104): 5 log10 norms */
int or_mg_residual(const or_state* o, double* residual);

/* overset donor table of the fine level (src/global_osg.F90:18-28): hosts,
 * nodes 3 ints (1-based i,j,k) and coef 3 doubles per donor; apply_mom = 1
 * interpolates q(1:4) after every fine smoother pass (the reference does
 * when nzone > 1), nu~ is interpolated after every des_eddy.  n = 0 clears. */
int or_set_overset(or_state* o, long n, const int* hosts, const double* coef, const int* nodes, int apply_mom);
/* one mg_bintp_mom (which = 0) or mg_bintp_tur (which = 1) on the stored state */
int or_bintp(or_state* o, int which);

/* one physical step of nz zones (one rank each in the reference) coupled by a
 * cross-zone donor table: hosts, nodes 4 ints per donor (i, j, k, zone;
 * 1-based), coef 3 doubles; nu~ is interpolated after every des_eddy, q(1:4)
 * after every fine pass when apply_mom.  Norms: sums over all zones divided
 * by zone 1's interior count (src/mg_residual.F90:36,71-82). */
int or_group_mg_driver(or_state* const* zs, int nz, long nd, const int* hosts, const double* coef,
                       const int* nodes, int apply_mom, int* itc, double* resid);
/* one cross-zone interpolation (which = 0 q(1:4), 1 nu~) with such a table */
int or_group_bintp(or_state* const* zs, int nz, long nd, const int* hosts, const double* coef,
                   const int* nodes, int which);

/* actuator-line model tables (desalm_turbine / desalm_polar of desalm.h;
 * copied); n_tur = 0 removes them */
int or_set_turbines(or_state* o, int n_tur, const desalm_turbine* tur, int n_polar, const desalm_polar* pol,
                    double length_scale, double vel_infty);

/* LAPACK SGTSV algorithm in FP64 (src/sgtsv.f:1-263); returns INFO */
int or_gtsv(int n, int nrhs, double* dl, double* d, double* du, double* b, int ldb);
/* The two case hooks the reference's mg_driver calls on every run:
 * conservation (src/conservation.F90, src/mg_driver.F90:64-75) after the
 * time shift, and wf_eta on both j walls after every fine-level solver_daf
 * (src/wf_mpi.F90:246-413, src/mg_driver.F90:218-224).  wf_newton = 1
 * solves for u_tau with the reference's own Newton chain (start 0.03 on the
 * first node, then each node from the previous one's result, stopped at
 * |f| < 1e-4); 0 converges every node's root to rounding, as the HIP
 * library does.  Hooks start off. */
#define OR_HOOK_CONSERVATION 1
#define OR_HOOK_WALL_FUNCTIONS 2
int or_set_hooks(or_state* o, int hooks, int wf_newton);
/* u_tau of one node: the converged root, and the reference's Newton from *x */
double or_wf_root(double u0, double yp0, double re0);
int or_wf_newt(double* x, double u0, double yp0, double re0);

/* row interchanges performed by or_gtsv so far (reset to 0 if reset != 0) */
long or_gtsv_pivots(int reset);

#ifdef __cplusplus
}
#endif
#endif
