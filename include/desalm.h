/*
 * desalm.h -- C ABI of libdesalm.so, the MI355X (gfx950) implementation of
 * the DES-ALM pseudo-time inner iteration: residual assembly, DAF implicit
 * operator with i/j/k tridiagonal line sweeps, SA/DES eddy-viscosity update
 * and the FAS multigrid V-cycle.
 *
 * Plain pointers and sizes only.  All floating-point arrays are double
 * precision and use the reference's own array-of-structures layout, i.e. a
 * field q(1:m, l) is m consecutive doubles per node and the node index l runs
 * i fastest, then j, then k, then multigrid level (src/init.F90:880-980 of
 * the reference, "le_idx" numbering on a single rank without ghost layers).
 * The library converts to its device structure-of-arrays layout internally.
 *
 * Return codes: 0 = success, < 0 = argument or HIP error (text available from
 * desalm_last_error; DESALM_ERR_NONFINITE for update norms that are NaN or
 * Inf), > 0 = LAPACK-INFO-like zero pivot reported by a line solve (1-based
 * row, in line order, of the first row whose pivot is zero or not finite).
 *
 * One context drives one HIP device on one stream; a context is not
 * re-entrant (the reference runs one MPI rank per context).
 */
#ifndef DESALM_H
#define DESALM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DESALM_MAX_LEVELS 6

/* desalm_mg_driver / desalm_group_mg_driver / desalm_step: the L1 update
 * norms of a pseudo-iteration are NaN or Inf.  The call stops after that
 * iteration; itc_out and the norm history hold it. */
#define DESALM_ERR_NONFINITE (-2)

/* Arithmetic used by the DAF factor and its line solves.  The residual,
 * metrics and multigrid restriction are always FP64. */
#define DESALM_PREC_FP64 0   /* plain FP64 everywhere */
#define DESALM_PREC_MIXED 1  /* FP32 DAF factor + sweeps as a defect correction */

/* Boundary types, as the reference's btype (ind3dmg.dat notes; src/bcond_fm.F90). */
#define DESALM_BC_INTERFACE 0
#define DESALM_BC_WALL 1
#define DESALM_BC_SYMMETRY 2
#define DESALM_BC_FREESTREAM 3
#define DESALM_BC_INFLOW 4
#define DESALM_BC_OUTLET_MOC 5   /* characteristic outlet (mg_brhs + nonreflect_ibc on i faces) */
#define DESALM_BC_PERIODIC 6       /* the reference's plane copy (the planes freeze) */
#define DESALM_BC_PERIODIC_WRAP 7  /* no counterpart in the reference: both planes of the
                                     * direction are images of the opposite interior planes
                                     * (q(1) = q(n-1), q(n) = q(2)) on the fine level, for q
                                     * and nu~; the faces of a direction are 7 together */
#define DESALM_BC_FREE_SURFACE 8

/*
 * Run parameters.  Mirrors the module state the reference reads inside
 * mg_driver (src/global.F90, src/global_param.F90, src/global_app.F90) as it
 * is filled by init_input_file (src/init.F90:176-228) and init_daf_bc
 * (src/init.F90:486-508) for a single zone on a single rank.
 * Indices in this struct (iter[], btype[], ep[][]) are 0-based versions of
 * the reference's 1-based arrays.
 */
/* Local pseudo time step of the DAF smoother (desalm_params.dtau_formula).
 * LIVE: rhs_daf_dtau (src/rhs_daf_dtau.F90:51-83), the routine solver_daf
 *       calls: the viscous limits use 1/Re + nu_t.
 * MET:  met_daf_dtau (src/met_daf_dtau.F90:53-93), the reference's other
 *       formula, which no reference path calls: the viscous limits use
 *       vnn1 Re (laminar viscosity only). */
#define DESALM_DTAU_LIVE 0
#define DESALM_DTAU_MET 1

typedef struct desalm_params {
    int32_t ng;                          /* multigrid levels (global.F90:17) */
    int32_t img, jmg, kmg;               /* fine-level node counts */
    int32_t icrs, jcrs, kcrs;            /* 1 = full coarsening in that direction */
    int32_t iter[DESALM_MAX_LEVELS];     /* smoother passes per level, iter(n) */
    int32_t itm;                         /* max pseudo-iterations, itm(1) */
    int32_t it_min;                      /* no convergence exit while itc <= it_min */
    double er_min;                       /* exit when max drop of log10 u,v,w norms < er_min */
    double eo_min;                       /* exit when max log10 u,v,w norm < eo_min */
    double ren;                          /* Reynolds number */
    double beta;                         /* artificial compressibility */
    double delti;                        /* physical time step */
    int32_t unsteady;                    /* 1: BDF2 term on (e_source = 1) */
    int32_t des;                         /* 1: SA/DES update of nu~ on the fine level */
    int32_t itk;                         /* SA sub-iterations per smoother pass */
    int32_t precision;                   /* DESALM_PREC_* */
    double cfl1, vnn1;                   /* DAF local time step limits */
    double cfl2, vnn2;                   /* SA local time step limits (dtev) */
    double pdiss_coef;                   /* pressure dissipation coefficient */
    double ep[2][3];                     /* ep(1,d): DAF implicit smoothing, ep(2,d): prolongation smoothing */
    int32_t btype[6];                    /* faces: i=1, i=max, j=1, j=max, k=1, k=max */
    int32_t bdir[6];                     /* symmetry-plane normal direction (1,2,3) */
    int32_t ifix, jfix, kfix;            /* 1-based node whose pressure is pinned to 0 */
    int32_t dtau_formula;                /* DESALM_DTAU_*: local pseudo time step formula */
    int32_t reserved[6];
    double body_force[3];                /* constant body force per unit mass in x, y, z, added to
                                          * the momentum equations of every level (-f / J in rh);
                                          * no counterpart in the reference (a channel's mean
                                          * pressure gradient); zero in the reference's cases */
} desalm_params;

typedef struct desalm_ctx desalm_ctx;

/* Allocate device state for all levels (replaces the module setup done by
 * init_mg, src/init.F90:1025-1124, and init_daf_bc, src/init.F90:439-511).
 * device: HIP device ordinal. */
int desalm_init(const desalm_params* p, int device, desalm_ctx** ctx);

/* Total node count over all levels and the offset (in nodes) of level n
 * (1-based), i.e. le_idx_b(ng) and le_idx_a(n)-1 of the reference. */
int64_t desalm_nodes_all(const desalm_ctx* ctx);
int64_t desalm_level_offset(const desalm_ctx* ctx, int n);

/* Fine-level grid coordinates x,y,z (img*jmg*kmg each).  Injects them to
 * the coarse levels (src/mg_inject.F90:2-64 as called from init) and
 * computes csi, eta, zet, aj on every level and dtev on the fine level
 * (src/mg_metrics.F90:56-300) on the device. */
int desalm_set_grid(desalm_ctx* ctx, const double* x, const double* y, const double* z);

/* Alternative to desalm_set_grid: metrics of all levels in the reference
 * layout csi(1:3, nall), eta(1:3, nall), zet(1:3, nall), aj(nall). */
int desalm_set_metrics(desalm_ctx* ctx, const double* csi, const double* eta,
                       const double* zet, const double* aj);

/* Fine-level coordinates only, for a caller that passes its own metrics with
 * desalm_set_metrics (as the Fortran drop-in does): the case hooks of
 * desalm_set_hooks read them.  Metrics are left as they are. */
int desalm_set_coordinates(desalm_ctx* ctx, const double* x, const double* y, const double* z);

/* Wall distance of the fine level (img*jmg*kmg), pre-scaled for DES as the
 * reference reads it from its grid file (src/init.F90:2313).  dtev may be
 * NULL (keep the value computed by desalm_set_grid) or a fine-level array. */
int desalm_set_wall_distance(desalm_ctx* ctx, const double* wd, const double* dtev);

/* Fine-level solution in the reference layout: q(1:5, n1), qn(1:5, n1),
 * qnm1(1:5, n1), xnut(n1), n1 = img*jmg*kmg.  Coarse levels are produced by
 * the V-cycle's own injections.  Any pointer may be NULL to leave that
 * field unchanged (get) or zero (set). */
int desalm_set_state(desalm_ctx* ctx, const double* q, const double* qn,
                     const double* qnm1, const double* xnut);
int desalm_get_state(desalm_ctx* ctx, double* q, double* qn, double* qnm1, double* xnut);

/*
 * One physical step's pseudo-time inner iteration: replaces
 * `subroutine mg_driver` (src/mg_driver.F90:2-323).  Shifts qnm1 <- qn <- q,
 * runs up to itm FAS V-cycles of DAF smoothing (SA/DES update on the fine
 * level when des = 1), applies the fine-grid BCs and stops on the
 * reference's convergence test (mg_ck_conver, src/mg_driver.F90:335-405).
 * itc_out: pseudo-iterations done.  resid_out (may be NULL): itm*5 doubles,
 * row it holds the log10 L1 update norms of mg_residual
 * (src/mg_residual.F90:2-104) after iteration it.
 * State stays on the device; use desalm_get_state to read it back.
 */
int desalm_mg_driver(desalm_ctx* ctx, int* itc_out, double* resid_out);

/*
 * One DAF smoother pass on level n (1-based) with host arrays of that level:
 * replaces `subroutine solver_daf` (src/solver_daf.F90:2-16,80-89).
 * q, qn, qnm1, pk, rh: (1:4, nl) where nl = nodes of level n; csi, eta, zet:
 * (1:3, nl); aj, xnut: (nl).  q, pk and rh are updated in place exactly as
 * the reference's intent(inout/out) dummies.  Copies in and out per call:
 * this is the parity entry, not the production path.
 */
int desalm_solver_daf(desalm_ctx* ctx, int n, int decide_recalc_rh, int decide_calc_pk,
                      double* q, const double* qn, const double* qnm1,
                      const double* csi, const double* eta, const double* zet,
                      const double* aj, const double* xnut, double* pk, double* rh);

/*
 * SA/DES eddy-viscosity update on the fine level with host arrays: replaces
 * `subroutine des_eddy` (src/des_eddy.F90:2-14).  q: (1:5, n1) updated in
 * place (component 5 = nu~); qn5, qnm1_5: nu~ of the two previous physical
 * steps (n1); csi, eta, zet (1:3, n1); aj, wd, dtev (n1); xnut (n1) output.
 */
int desalm_des_eddy(desalm_ctx* ctx, double* q, const double* qn5, const double* qnm1_5,
                    const double* csi, const double* eta, const double* zet,
                    const double* aj, const double* wd, const double* dtev, double* xnut);

/*
 * Batched tridiagonal solves with the reference's LAPACK SGTSV semantics
 * (src/sgtsv.f:1-263, Gaussian elimination with partial pivoting) in FP64 on
 * the device, one system per column: nsys systems of size n stored
 * column-major (entry r of system s at [s*n + r]); dl and du hold n-1 used
 * entries per system at [s*n + 0 .. n-2].  All arrays are host arrays and
 * are overwritten (b with the solution).  info_out (may be NULL): nsys ints,
 * the SGTSV INFO per system.
 */
int desalm_gtsv_batch(int n, int nsys, double* dl, double* d, double* du, double* b,
                      int* info_out);

/*
 * desalm_gtsv_batch without any device allocation: no counterpart in the
 * reference.  For callers that cannot allocate device memory (serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, kernel and results as
 * desalm_gtsv_batch on the current HIP device; the device copies of the four
 * arrays and INFO live in one fixed device array of the library (9 MiB
 * per GPU, held until the process ends), so n * nsys may not exceed
 * DESALM_GTSV_NOALLOC_MAX (larger batches return < 0 with a message).  All
 * calls on one GPU share that array: calls that overlap on one GPU race.
 */
#define DESALM_GTSV_NOALLOC_MAX 262144
int desalm_gtsv_batch_noalloc(int n, int nsys, double* dl, double* d, double* du, double* b,
                              int* info_out);

/*
 * Overset donor table of the fine level, in place of the tables that
 * init_osg_interface reads (src/global_osg.F90:18-28, src/init.F90:1979-2000),
 * for donors and fringe nodes of this one zone: donor d interpolates
 * trilinearly from the cell whose lower corner is hosts[3d..3d+2] (1-based
 * i, j, k) with weights coef[3d..3d+2] into the fine node nodes[3d..3d+2].
 * desalm_mg_driver then interpolates nu~ (and sets xnut) after every des_eddy
 * (mg_bintp_tur, src/mg_driver.F90:182), and q(1:4) after every fine smoother
 * pass when apply_mom != 0 (mg_bintp_mom, src/mg_driver.F90:241-244; the
 * reference runs it when nzone > 1).  All donor values are formed before any
 * fringe node is written.  Fringe nodes must be distinct.  n = 0 removes the
 * table.
 */
int desalm_set_overset(desalm_ctx* ctx, int64_t n, const int32_t* hosts, const double* coef,
                       const int32_t* nodes, int apply_mom);

/* One interpolation with the stored table on the device state: which = 0
 * mg_bintp_mom (src/mg_bintp_mom.F90:2-172), 1 mg_bintp_tur
 * (src/mg_bintp_tur.F90:2-182). */
int desalm_bintp(desalm_ctx* ctx, int which);

/*
 * Several zones coupled by overset interpolation: the reference's nzone > 1
 * runs, in which every zone has its own ranks and the zones meet only in
 * mg_bintp_tur / mg_bintp_mom (src/mg_driver.F90:182,241-244) and in the norm
 * reduction of mg_residual (src/mg_residual.F90:71-82).  A group borrows nz
 * contexts (one per zone, zone 1 first, all on one device, with the same
 * multigrid levels, iter, itm, it_min, er_min, eo_min, des and itk, which the
 * reference holds globally) and runs their V-cycles in lockstep on zone 1's
 * stream.  The contexts stay owned by the caller and keep their state; a
 * zone's own single-zone table (desalm_set_overset) is not used by the group.
 */
typedef struct desalm_group desalm_group;
int desalm_group_create(desalm_ctx* const* zones, int nz, desalm_group** group);

/* Cross-zone donor table (src/global_osg.F90:18-28): hosts[4d..4d+3] = lower
 * corner i, j, k and zone (1-based) of donor d's host cell, coef[3d..3d+2]
 * its trilinear weights, nodes[4d..4d+3] = i, j, k and zone of the fringe
 * node it fills.  nu~ is interpolated after every des_eddy, q(1:4) after
 * every fine smoother pass when apply_mom != 0.  All donor values are formed
 * before any fringe node is written; fringe nodes must be distinct within a
 * zone.  n = 0 removes the table. */
int desalm_group_set_overset(desalm_group* group, int64_t n, const int32_t* hosts, const double* coef,
                             const int32_t* nodes, int apply_mom);

/* One physical step of every zone (src/mg_driver.F90:2-323 on all ranks):
 * itc_out and resid_out as for desalm_mg_driver, with the norms summed over
 * all zones and divided by zone 1's interior node count, as the reference's
 * root rank does (src/mg_residual.F90:36,71-82). */
int desalm_group_mg_driver(desalm_group* group, int* itc_out, double* resid_out);
int desalm_group_destroy(desalm_group* group);
/* One launch per kernel over all zones of the group (on = 1, the default):
 * every kernel of a pseudo-iteration runs once over the workgroups of all
 * zones, each workgroup finding its zone in a block-descriptor table of the
 * zones' levels, fields and constants.  The zones' characteristic-outlet
 * rows, actuator-line forcing and hub zeroing, the wall functions of zone 1
 * and the donor exchanges keep their own launches.  on = 0 launches zone by
 * zone.  Zones of mixed precision forms, or more than 16 zones, always run
 * zone by zone.  No counterpart in the reference, whose zones run on their
 * own MPI ranks (mg_driver.F90:241-247). */
int desalm_group_set_batched(desalm_group* group, int on);

/*
 * Actuator-line model (rhs_lines, src/rhs_lines.F90:47-483, called from
 * solver_daf at src/solver_daf.F90:378-381; hub zeroing in bcond_fm,
 * src/bcond_fm.F90:328-355): one rotor as the tables that init_lines builds
 * (src/init.F90:1156-1710, module global_turb).  Node indices are 1-based
 * on the fine level.  Arrays follow the Fortran order of the reference's
 * arrays for one turbine:
 *   node_indexes[(elem-1) + nelems*((nang-1) + nangles*(c-1))], c = 1..3
 *   node_indexes_pos[same with c = 1..27]: element x, y, z, then the corners
 *     000, 100, 110, 010, 001, 101, 111, 011 (x, y, z each)
 *   disk_dat[(j-1) + jmg*((k-1) + kmg*(c-1))], c = 1..8
 * polar_before / polar_after[(j-1) + jmg*(k-1)]: 0-based index into the
 * polars passed with the turbines, for the element at node (., j, k).
 */
typedef struct desalm_turbine {
    int32_t nblades, nelems, nangles;
    int32_t itur1, itur2;                 /* i range zeroed in u at hub nodes */
    double tur_rad, hub_rad, tsr;         /* radii in length_scale units; tip speed ratio */
    double angle;                         /* angle of the reference blade (rad) */
    double dtheta_interp;                 /* angle bin of node_indexes (rad) */
    const int32_t* node_indexes;
    const double* node_indexes_pos;
    const double* disk_dat;
    const int32_t* polar_before;
    const int32_t* polar_after;
    int32_t n_disk;                       /* hub/disk nodes: disk_node_coordinates */
    const int32_t* disk_nodes;            /* [n_disk*3] i, j, k */
    int32_t n_ker;                        /* spreading nodes: disk_node_coordinates_ker */
    const int32_t* ker_nodes;             /* [n_ker*3] i, j, k */
    const double* ker_pos;                /* [n_ker*3] x, y, z */
    const double* ker_dat;                /* [n_ker*3] disk_dat_ker components 1, 2 and 4 at those nodes */
} desalm_turbine;

/* An airfoil polar: n points of (angle of attack in degrees, ascending, Cl, Cd),
 * the content of one airfoil file that rhs_lines reads (src/rhs_lines.F90:232-300). */
typedef struct desalm_polar {
    int32_t n;
    const double* angle;
    const double* cl;
    const double* cd;
} desalm_polar;

/* Turn the actuator-line forcing on with n_tur rotors, their polars and the
 * scales of global_turb (length_scale, vel_infty); n_tur = 0 turns it off.
 * The forcing is added to rh(2:4) in every smoother pass on every level
 * (not in the recalc branch, src/solver_daf.F90:432-484), and bcond_fm zeroes
 * u on the hub nodes.  An angle of attack outside a polar takes the polar's
 * end value (the reference leaves the previous element's value there). */
int desalm_set_turbines(desalm_ctx* ctx, int32_t n_tur, const desalm_turbine* turbines, int32_t n_polar,
                        const desalm_polar* polars, double length_scale, double vel_infty);

/* Read a device field of level n back in the reference layout.
 * which: 0 q(5), 1 qn(5), 2 qnm1(5), 3 qold_mg(5), 4 rh(4), 5 pk(4),
 * 6 xnut(1), 7 csi(3), 8 eta(3), 9 zet(3), 10 aj(1), 11 dtev(1), 12 wd(1).
 * out must hold ncomp * nodes(level n) doubles. */
int desalm_get_level_field(desalm_ctx* ctx, int which, int n, double* out);

/* Device timing of the last desalm_mg_driver call (milliseconds, from HIP
 * events on the context's stream). */
double desalm_last_driver_ms(const desalm_ctx* ctx);

/* Rows, summed over every line solve of the last desalm_mg_driver /
 * desalm_solver_daf / desalm_des_eddy call, at which LAPACK SGTSV would have
 * interchanged rows (|d_i| < |dl_i|).  The line solves eliminate without
 * interchanges; a non-zero count says their row order departed from the
 * reference's SGTSV on that call. */
int64_t desalm_last_pivot_events(const desalm_ctx* ctx);

/* Failure detection of the last desalm_mg_driver / desalm_step /
 * desalm_solver_daf call (no counterpart in the reference, which prints and
 * carries on): zero_pivot_row = the INFO of the line solves (first row,
 * 1-based in line order, whose pivot was zero or not finite; 0 = none),
 * nonfinite_iteration = the pseudo-iteration whose update norms were NaN or
 * Inf (0 = none), diverged_iteration = the first pseudo-iteration whose
 * largest u, v, w update norm exceeded the first iteration's by more than
 * three decades (0 = none).  fallbacks: physical steps so far that the
 * FP32-DAF form or a warm start did not carry (a zero pivot, non-finite or
 * diverging norms) and that were redone from the step's entry state with
 * the plain FP64 method and the reference's start; the status then describes
 * the plain run.  Any pointer may be NULL. */
int desalm_last_status(const desalm_ctx* ctx, int32_t* zero_pivot_row, int32_t* nonfinite_iteration,
                       int32_t* diverged_iteration, int64_t* fallbacks);

/* Device time per kernel family and level, from HIP events recorded on the
 * context's stream around each launch inside desalm_mg_driver.
 * desalm_profile(ctx, 1) clears the counters and starts recording, 0 stops.
 * kernel: 0 .. desalm_profile_name() returns NULL; level 1-based. */
int desalm_profile(desalm_ctx* ctx, int enable);
int desalm_profile_get(const desalm_ctx* ctx, int kernel, int level, int64_t* count, double* total_ms);
const char* desalm_profile_name(int kernel);

/* Device memory the context holds: fields, scratch, rotor tables and the
 * warm start (bytes; overset tables of a group excluded). */
int64_t desalm_device_bytes(const desalm_ctx* ctx);

/* Launch shape of the LDS-tiled stencil kernels: k planes per workgroup of
 * the viscous kernel on each level (kc_visc[0 .. ng-1]) and of the SA
 * right-hand side (fine level); 0 picks the default (about 2048 workgroups).
 * The shape changes results only by rounding (where a workgroup's first k
 * face is evaluated, multiply-adds may be contracted differently).
 * desalm_autotune times every candidate
 * (1, 2, 3, 4, 6, 8, 12, 16 planes) over reps launches on the context's
 * current state and keeps the fastest; its launches only rewrite scratch.
 * desalm_get_tiles reports the shapes in use.  No counterpart in the
 * reference. */
int desalm_set_tiles(desalm_ctx* ctx, const int* kc_visc, int kc_sa);
int desalm_get_tiles(const desalm_ctx* ctx, int* kc_visc, int* kc_sa);
int desalm_autotune(desalm_ctx* ctx, int reps);

/* Convergence controls of the following desalm_mg_driver calls: at most
 * itm pseudo-iterations, stop after it_min when the largest u, v, w log10
 * norm has dropped below er_min relative to the first iteration or below
 * eo_min absolutely (mg_ck_conver, mg_driver.F90:372-398; the reference reads
 * them from its input deck, init.F90:176-228).  log_resid_out of a later
 * call must hold 5 * itm doubles. */
int desalm_set_convergence(desalm_ctx* ctx, int itm, int it_min, double er_min, double eo_min);

/* Random-flow-generator inflow (rfg_inlet.F90:107-173 with rfg.F90:329-374,
 * the method of Smirnov, Shi & Celik, J. Fluids Eng. 123 (2001) 359-371):
 * before each desalm_step's V-cycles, at time = time0 + (steps since this
 * call) * delti, every interior node (j, k) of the plane i = 1 gets
 *   (u, v, w) = um + M (c o v1),
 *   v1 = sqrt(2 / N) sum_n [p_n cos(theta_n) + q_n sin(theta_n)],
 *   theta_n = sum_d k_nd (ls / ts) / c_d x_d / ls + omega_n time / ts,
 * with M the eigenvectors (columns) and c the square roots of the eigenvalues
 * of the node's Reynolds-stress tensor.  The caller supplies the N modes
 * (k_n, omega_n normal; p_n = zeta_n x k_n, q_n = xi_n x k_n; the reference
 * draws them with its own generator) and per (j, k), index j + jmg k: um
 * (3), M (9, column-major), c (3), ts, ls.  n_modes = 0 removes it. */
int desalm_set_rfg(desalm_ctx* ctx, int32_t n_modes, const double* k, const double* omega, const double* p,
                   const double* q, const double* um, const double* m, const double* c, const double* ts,
                   const double* ls, double time0);

/* ---- Physical time step (main.F90:38-62) ----
 * One step of the physical-time loop around mg_driver: every rotor's blade
 * angle advances (angle > 2 pi wraps by -2 pi first, then angle += dtheta with
 * dtheta = TSR delti length_scale / tur_rad, main.F90:42-49 and
 * init.F90:1447), the random-flow-generator inflow is applied when set
 * (main.F90:58), then desalm_mg_driver runs, from the warm start selected by
 * desalm_ws_mode. */
int desalm_step(desalm_ctx* ctx, int* itc_out, double* log_resid_out);

/* ---- Outputs of a physical step (main.F90:71-75) ----
 * log10_out[0..4]: the real residual of output_real_residual
 * (src/main.F90:191-300), log10 sqrt(sum over the fine interior of
 * (q - qn)^2 / interior count) per component of the current state: after a
 * desalm_step, the change over the step (the 'conver_real' line).  Uses the
 * norm buffer of the driver; not during a driver call. */
int desalm_real_residual(desalm_ctx* ctx, double* log10_out);
/* Monitor points of output_probe (src/main.F90:536-555): for n points
 * ijk[3p..3p+2] (1-based fine-level i, j, k), out[6p..6p+5] = q(1:5) and
 * xnut there (the 'history' line of each point). */
int desalm_probe(desalm_ctx* ctx, int32_t n, const int32_t* ijk, double* out);

/* ---- Learned warm start ----
 * No counterpart in the reference, which starts every physical step's
 * pseudo-iterations from the previous solution (mg_driver.F90:48-51).  With
 * a warm start, desalm_mg_driver replaces that start on the fine level's
 * interior, after the qn / qnm1 shift, by
 *   DESALM_WS_EXTRAP   2 q^n - q^(n-1)
 *   DESALM_WS_LEARNED  q^n + a d1 + b d2, d1 = q^n - q^(n-1),
 *                      d2 = q^(n-1) - q^(n-2), with (a - 1, b) per node and
 *                      per variable the 10 outputs of a multilayer
 *                      perceptron (95 inputs, two tanh layers of `hidden`
 *                      units, FP32).  Inputs: for each of p, u, v, w, nu~ the
 *                      18 differences d1, d2 at the node and its eight
 *                      neighbours (the six face neighbours and the second
 *                      neighbours along i, clamped to the block) divided by
 *                      their root mean square, then q^n at the node;
 *                      standardised by the statistics.
 * nu~ keeps its floor of 1e-7.  The pseudo-iterations then run to the same
 * tolerance as from the reference's start. */
#define DESALM_WS_NONE 0
#define DESALM_WS_EXTRAP 1
#define DESALM_WS_LEARNED 2

/* Allocate the warm start: network of `hidden` (1..64) units per layer with
 * seeded Xavier-uniform weights and a zero output layer (so it starts as
 * linear extrapolation), a sample buffer of `capacity` rows, and q^(n-2)
 * (set to qnm1 here and by desalm_set_state, shifted by every
 * desalm_mg_driver).  Resets any earlier warm start. */
int desalm_ws_configure(desalm_ctx* ctx, int hidden, int64_t samples_per_record, int64_t capacity,
                        uint64_t seed);
int desalm_ws_mode(desalm_ctx* ctx, int mode);
/* After a physical step: samples_per_record seeded interior nodes of the
 * current state give (features from q^n, q^(n-1), q^(n-2); targets r = q -
 * 2 q^n + q^(n-1), d1, d2) rows, appended to the buffer (the oldest are
 * overwritten when it is full). */
int desalm_ws_record(desalm_ctx* ctx);
/* iters Adam steps (lr, beta 0.9 / 0.999, eps 1e-8) on batches of `batch`
 * rows drawn from the buffer, minimising the mean over rows and variables of
 * ((a - 1) d1 + b d2 - r)^2 / scale^2.  The statistics (feature means and
 * 1/sd, and per variable the root mean square of d1 as scale) are fixed by
 * the first call.  loss_out, if not NULL, receives the loss of every step
 * (iters doubles). */
int desalm_ws_train(desalm_ctx* ctx, int iters, int batch, double lr, double* loss_out);
/* Packed parameters (count desalm_ws_param_count: W1 96 x hidden, W2
 * (hidden + 1) x hidden, W3 (hidden + 1) x 10, row-major, last row of each
 * the bias) and statistics (200 floats: 95 feature means, 95 feature 1/sd,
 * 5 target scales, 5 unused); get when `get` is not NULL, then set when
 * `set` is not NULL. */
int desalm_ws_param_count(const desalm_ctx* ctx);
int desalm_ws_params(desalm_ctx* ctx, float* get, const float* set);
int desalm_ws_stats(desalm_ctx* ctx, float* get, const float* set);
/* q^(n-2) of the fine level, (nodes, 5) as desalm_get_state: get when not
 * NULL, then set when not NULL (a saved state is restored with
 * desalm_set_state followed by this). */
int desalm_ws_history(desalm_ctx* ctx, double* get, const double* set);
/* Blade angle (rad) of rotor t, get and/or set (global_turb's angle(n_t)). */
int desalm_turbine_angle(desalm_ctx* ctx, int t, double* get, const double* set);
/* Unstandardised features (m x 95) and targets (m x 15: r, d1, d2 per
 * variable) of the current state at interior ordinals 0 .. m-1 (i fastest),
 * as desalm_ws_record builds them. */
int desalm_ws_features(desalm_ctx* ctx, int64_t m, float* x, float* t);
/* The network on host rows: y (m x 10) from standardised features x
 * (m x 95); and the loss against targets t (m x 15, already divided by the
 * scales) with its gradient with respect to the packed parameters. */
int desalm_mlp_forward(desalm_ctx* ctx, const float* x, int64_t m, float* y);
int desalm_mlp_loss_grad(desalm_ctx* ctx, const float* x, const float* t, int64_t m, float* grad, double* loss);

/*
 * The two case hooks the reference's mg_driver calls on every run of zone 1
 * on one rank (bit flags, 0 = none, the default):
 *  DESALM_HOOK_CONSERVATION: `conservation` (src/conservation.F90:1-174,
 *    called at src/mg_driver.F90:64-75) after the time shift: U/J summed
 *    over the i = 1 and i = img planes, U/J on the i = img plane scaled by
 *    fluxin / fluxout and u, v, w rebuilt there.  Divides by the outlet flux:
 *    a closed i = img face gives NaN, as in the reference.
 *  DESALM_HOOK_WALL_FUNCTIONS: `wf_eta` on both j walls after every
 *    fine-level DAF pass (src/wf_mpi.F90:246-413, src/mg_driver.F90:218-224):
 *    log-law velocity and nu~ at the first node off the wall from the
 *    friction velocity of the node two off.  u_tau is solved to rounding at
 *    every node; the reference's Newton chain stops at |f| < 1e-4, so the two
 *    differ by up to that tolerance (about 1e-4 of the velocity).
 * Needs desalm_set_grid or desalm_set_coordinates (the hooks use the grid
 * derivatives).  A group of
 * zones rejects hooks (the reference's flux sums run over all ranks).
 * A desalm_ws_mode warm start is applied before conservation.
 */
#define DESALM_HOOK_CONSERVATION 1
#define DESALM_HOOK_WALL_FUNCTIONS 2
int desalm_set_hooks(desalm_ctx* ctx, int hooks);

/*
 * ---- The reference's files (host only: no device, no context) ----
 *
 * grid and solu are Fortran unformatted sequential files (every record framed
 * by 4-byte byte counts; records over 2^31 - 9 bytes split into subrecords as
 * gfortran writes them).  real_bytes: 4 (the reference's REAL*4,
 * src/precision.F90:11), 8 (an FP64 build), or 0 on reading (taken from the
 * first record).  dims[3 * nz + 0..2] = img, jmg, kmg of zone nz; the arrays of
 * the zones follow each other, nodes i fastest, then j, then k.  Errors are
 * reported by desalm_last_error(NULL).
 */

/* The grid file `grid` (src/init.F90:2226-2241): x, y, z of every zone, and
 * the wall distance wd when turbulence != 0. */
int desalm_read_grid_file(const char* path, int32_t nzone, const int32_t* dims, int32_t turbulence,
                          int32_t real_bytes, double* x, double* y, double* z, double* wd);

/* The solution file `solu` (src/init.F90:2384-2460) into q(1:5, n) and
 * qn(1:5, n) (the layout of desalm_set_state) and xnut(n) when turbulence.
 * Without turbulence the reference reads one record per field; the
 * one-record-per-component files of the restart output are read as well. */
int desalm_read_solu_file(const char* path, int32_t nzone, const int32_t* dims, int32_t turbulence,
                          int32_t real_bytes, double* q, double* qn, double* xnut);

/* The restart output of one rank (src/main.F90:343-372, solufiles/solu.NNN):
 * five records of q, xnut when turbulence, five of qn, each of n1 reals,
 * rounded to REAL*4 when real_bytes = 4.  Read back by
 * desalm_read_solu_file with nzone = 1. */
int desalm_write_solu_file(const char* path, int64_t n1, int32_t turbulence, int32_t real_bytes,
                           const double* q, const double* qn, const double* xnut);

/* `filestat` (src/init.F90:266-277): "time0 nt1" and one blade angle per
 * rotor, list-directed; an angle >= 2 pi is reduced by 2 pi as the reference
 * does on reading.  The writer follows src/main.F90:379-386. */
int desalm_read_filestat(const char* path, int32_t n_tur, double* time0, int32_t* nt1, double* angle);
int desalm_write_filestat(const char* path, double time, int32_t ntime, int32_t n_tur, const double* angle);

/* The input deck ind3dmg.dat (src/init.F90:176-228), list-directed.  ng, the
 * number of multigrid levels, is a compile-time constant of the reference
 * (src/global.F90:17) and an argument here.  params receives every run
 * parameter the deck holds (itm = itm(1)); unsteady and des take the
 * reference's defaults (src/global_param.F90:82, :121) and precision FP64.
 * The single zone of the reference (nzone = 1, src/global.F90:20) is read. */
#define DESALM_DECK_MAX_TURBINES 32
typedef struct desalm_deck_turbine {
    int32_t itur1, itur2, jtur1, jtur2, ktur1, ktur2;  /* 1-based node box of the rotor */
    double tur_rad, hub_rad, tsr;
    int32_t nblades, nelems;
} desalm_deck_turbine;

typedef struct desalm_deck {
    desalm_params params;
    int32_t ns, icycle, nt2, icn, icnw, irk, checkpoint, nzfix;
    int32_t itm[DESALM_MAX_LEVELS];      /* itm(1:ng) */
    double alfa[4], eps[4];
    double cdes, ckin, cein;
    int32_t monitor_num_points;
    int32_t monitor_point_ijk[10][4];    /* point j: i, j, k, zone */
    double length_scale, vel_infty;
    int32_t turbines, node_indexes_load; /* logicals: 0 or 1 */
    int32_t n_tur;
    desalm_deck_turbine turbine[DESALM_DECK_MAX_TURBINES];
    int32_t nzblanking;                  /* 0 or 1 (nba_max = 1, src/global.F90:21) */
    int32_t blktype[6];
    int32_t blanking[3][2];              /* blanking(d, 1:2): index range in direction d */
    int32_t reserved[8];
} desalm_deck;

int desalm_read_deck(const char* path, int32_t ng, desalm_deck* deck);

const char* desalm_last_error(const desalm_ctx* ctx);
int desalm_finalize(desalm_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif /* DESALM_H */
