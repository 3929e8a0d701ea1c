/*
 * desalm_oracle.c -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
 *
 * CPU FP64 restatement of the reference's pseudo-time inner iteration
 * (jcsandov/DES-ALM, Fortran 90).  It is the checker of the HIP path: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 * Loops, operand order and quirks follow the reference routine cited above
 * each function; single zone, single rank (every MPI neighbour is
 * MPI_PROC_NULL, no ghost layers, src/init.F90:575,671-680), no blanking
 * boxes (nblk = 0).  The case hooks are options of a run (or_set_hooks:
 * conservation, and the wall functions with the reference's own Newton
 * chain or an exact u_tau), as are the actuator-line forcing with the hub
 * zeroing (or_set_turbines), inflow faces (btype 4) and the characteristic
 * outlet (btype 5); several zones run through the group driver.
 *
 * Pinning of each part (what checks it beyond this restatement) is listed in
 * the tests that import it: the reference Fortran compiled from its own
 * sources (oracle/ref, all composed routines), LAPACK dgtsv via scipy
 * (or_gtsv), closed forms and invariants (metrics, residual terms, transfer
 * operators).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "desalm_oracle.h"

#define NGM (DESALM_MAX_LEVELS + 1)

struct or_state {
    desalm_params p;
    int ng;
    int img[NGM], jmg[NGM], kmg[NGM];
    long off[NGM];
    long nl[NGM];
    long nall;
    double dc[NGM], de[NGM], dz[NGM];
    double e_source;
    double cvin;
    double sa[6][5], sb[6][5]; /* sa(m,b) -> sa[b-1][m-1] */
    double *x, *y, *z;
    double *xc, *xe, *xz, *yc, *ye, *yz, *zc, *ze, *zz;
    double *csi, *eta, *zet, *aj;
    double *q, *qn, *qnm1, *qold_mg, *rh, *pk;
    double *xnut, *wd, *dtev;
    double init_resd[5];
    int itc;
    /* overset donor table of one zone on one rank (global_osg.F90:18-28):
     * donor d interpolates from the cell with lower corner hosts(1:3,d) at
     * trilinear weights coef(1:3,d) into the fine node nodes(1:3,d) */
    long nover;
    int* ohosts;    /* 3 per donor, 1-based i,j,k */
    double* ocoef;  /* 3 per donor */
    int* onodes;    /* 3 per donor, 1-based i,j,k */
    int apply_mom;  /* mg_bintp_mom runs (the reference's nzone > 1) */
    /* case hooks of mg_driver (OR_HOOK_*): the outlet mass-flux rescale and
     * the j-wall wall functions; wf_newton selects the reference's own
     * Newton chain (1) or a root converged to rounding (0) */
    int hooks, wf_newton;
    /* grid ratio of the inflow pressure extrapolation per face (module
     * global_app's rat, init_outlet_bc, src/init.F90:2750-2891); NULL where
     * the face is not an inflow */
    double* rat[6];
    /* actuator-line model (module global_turb): deep copies of the tables */
    int n_tur, n_polar;
    desalm_turbine* tur;
    desalm_polar* pol;
    double length_scale, vel_infty;
};

/* constants of src/precision.F90 */
static const double one_half = 0.5, pt5 = 0.5, pt25 = 0.25, one_fourth = 0.25;
static const double onept5 = 1.5, one_pt_five = 1.5;

#define LEVEL(n)                                   \
    const int ni = o->img[n], nj = o->jmg[n], nk = o->kmg[n]; \
    const long off = o->off[n];                    \
    (void)off
#define IX(i, j, k) ((long)((i)-1) + (long)ni * ((long)((j)-1) + (long)nj * (long)((k)-1)))

static double dmin2(double a, double b) { return a < b ? a : b; }
static double dmax2(double a, double b) { return a > b ? a : b; }

/* ------------------------------------------------------------------ */
/* LAPACK SGTSV (vendored LAPACK 3.0, src/sgtsv.f:89-259), FP64.        */
/* dl(1..n-1), d(1..n), du(1..n-1), b(ldb, nrhs); 0-based storage.     */
/* ------------------------------------------------------------------ */
static long g_pivots = 0; /* row interchanges done by or_gtsv since the last reset */

long or_gtsv_pivots(int reset) {
    long v = g_pivots;
    if (reset) g_pivots = 0;
    return v;
}

int or_gtsv(int n, int nrhs, double* dl, double* d, double* du, double* b, int ldb) {
    int i, jr;
    double fact, temp;
    if (n < 0) return -1;
    if (nrhs < 0) return -2;
    if (ldb < (n > 1 ? n : 1)) return -7;
    if (n == 0) return 0;
    /* forward elimination with row interchanges (sgtsv.f:108-192) */
    for (i = 0; i < n - 2; i++) {
        if (fabs(d[i]) >= fabs(dl[i])) {
            if (d[i] != 0.0) {
                fact = dl[i] / d[i];
                d[i + 1] = d[i + 1] - fact * du[i];
                for (jr = 0; jr < nrhs; jr++) b[i + 1 + jr * ldb] = b[i + 1 + jr * ldb] - fact * b[i + jr * ldb];
            } else {
                return i + 1;
            }
            dl[i] = 0.0;
        } else {
            g_pivots++;
            fact = d[i] / dl[i];
            d[i] = dl[i];
            temp = d[i + 1];
            d[i + 1] = du[i] - fact * temp;
            dl[i] = du[i + 1];
            du[i + 1] = -fact * dl[i];
            du[i] = temp;
            for (jr = 0; jr < nrhs; jr++) {
                temp = b[i + jr * ldb];
                b[i + jr * ldb] = b[i + 1 + jr * ldb];
                b[i + 1 + jr * ldb] = temp - fact * b[i + 1 + jr * ldb];
            }
        }
    }
    if (n > 1) {
        i = n - 2;
        if (fabs(d[i]) >= fabs(dl[i])) {
            if (d[i] != 0.0) {
                fact = dl[i] / d[i];
                d[i + 1] = d[i + 1] - fact * du[i];
                for (jr = 0; jr < nrhs; jr++) b[i + 1 + jr * ldb] = b[i + 1 + jr * ldb] - fact * b[i + jr * ldb];
            } else {
                return i + 1;
            }
        } else {
            g_pivots++;
            fact = d[i] / dl[i];
            d[i] = dl[i];
            temp = d[i + 1];
            d[i + 1] = du[i] - fact * temp;
            du[i] = temp;
            for (jr = 0; jr < nrhs; jr++) {
                temp = b[i + jr * ldb];
                b[i + jr * ldb] = b[i + 1 + jr * ldb];
                b[i + 1 + jr * ldb] = temp - fact * b[i + 1 + jr * ldb];
            }
        }
    }
    if (d[n - 1] == 0.0) return n;
    /* back substitution (sgtsv.f:232-257) */
    for (jr = 0; jr < nrhs; jr++) {
        double* bb = b + jr * ldb;
        bb[n - 1] = bb[n - 1] / d[n - 1];
        if (n > 1) bb[n - 2] = (bb[n - 2] - du[n - 2] * bb[n - 1]) / d[n - 2];
        for (i = n - 3; i >= 0; i--) bb[i] = (bb[i] - du[i] * bb[i + 1] - dl[i] * bb[i + 2]) / d[i];
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* setup                                                               */
/* ------------------------------------------------------------------ */
static double* zalloc(long n) {
    double* p = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    return p;
}

or_state* or_create(const desalm_params* p) {
    int n, b;
    or_state* o;
    if (!p || p->ng < 1 || p->ng > DESALM_MAX_LEVELS) return NULL;
    for (b = 0; b < 6; b++)
        if (p->btype[b] < 0 || p->btype[b] > 8) return NULL;
    o = (or_state*)calloc(1, sizeof(or_state));
    o->p = *p;
    o->ng = p->ng;
    /* level sizes, init_input_file (src/init.F90:254-260) */
    o->img[1] = p->img;
    o->jmg[1] = p->jmg;
    o->kmg[1] = p->kmg;
    for (n = 2; n <= o->ng; n++) {
        o->img[n] = (o->img[n - 1] + p->icrs) / (1 << p->icrs);
        o->jmg[n] = (o->jmg[n - 1] + p->jcrs) / (1 << p->jcrs);
        o->kmg[n] = (o->kmg[n - 1] + p->kcrs) / (1 << p->kcrs);
    }
    o->nall = 0;
    for (n = 1; n <= o->ng; n++) {
        o->off[n] = o->nall;
        o->nl[n] = (long)o->img[n] * o->jmg[n] * o->kmg[n];
        o->nall += o->nl[n];
        /* init_mg (src/init.F90:1116-1120) */
        o->dc[n] = 1.0 / pow(2.0, (double)((n - 1) * p->icrs));
        o->de[n] = 1.0 / pow(2.0, (double)((n - 1) * p->jcrs));
        o->dz[n] = 1.0 / pow(2.0, (double)((n - 1) * p->kcrs));
    }
    o->e_source = p->unsteady ? 1.0 : 0.0;
    o->cvin = 0.1 / p->ren;
    /* init_daf_bc (src/init.F90:486-508) */
    for (b = 0; b < 6; b++) {
        int m;
        for (m = 0; m < 5; m++) {
            o->sa[b][m] = 4.0 / 3.0;
            o->sb[b][m] = -(1.0 / 3.0);
        }
        if (p->btype[b] == 1)
            for (m = 1; m < 4; m++) o->sa[b][m] = o->sb[b][m] = 0.0;
        if (p->btype[b] == 2 && p->bdir[b] >= 1 && p->bdir[b] <= 3)
            o->sa[b][p->bdir[b]] = o->sb[b][p->bdir[b]] = 0.0;
    }
    o->x = zalloc(o->nall); o->y = zalloc(o->nall); o->z = zalloc(o->nall);
    o->xc = zalloc(o->nall); o->xe = zalloc(o->nall); o->xz = zalloc(o->nall);
    o->yc = zalloc(o->nall); o->ye = zalloc(o->nall); o->yz = zalloc(o->nall);
    o->zc = zalloc(o->nall); o->ze = zalloc(o->nall); o->zz = zalloc(o->nall);
    o->csi = zalloc(3 * o->nall); o->eta = zalloc(3 * o->nall); o->zet = zalloc(3 * o->nall);
    o->aj = zalloc(o->nall);
    o->q = zalloc(5 * o->nall); o->qn = zalloc(5 * o->nall); o->qnm1 = zalloc(5 * o->nall);
    o->qold_mg = zalloc(5 * o->nall);
    o->rh = zalloc(4 * o->nall); o->pk = zalloc(4 * o->nall);
    o->xnut = zalloc(o->nall); o->wd = zalloc(o->nall); o->dtev = zalloc(o->nall);
    return o;
}

void or_destroy(or_state* o) {
    if (!o) return;
    free(o->x); free(o->y); free(o->z);
    free(o->xc); free(o->xe); free(o->xz); free(o->yc); free(o->ye); free(o->yz);
    free(o->zc); free(o->ze); free(o->zz);
    free(o->csi); free(o->eta); free(o->zet); free(o->aj);
    free(o->q); free(o->qn); free(o->qnm1); free(o->qold_mg); free(o->rh); free(o->pk);
    free(o->xnut); free(o->wd); free(o->dtev);
    free(o->ohosts); free(o->ocoef); free(o->onodes);
    for (int b = 0; b < 6; b++) free(o->rat[b]);
    or_set_turbines(o, 0, NULL, 0, NULL, 0.0, 0.0);
    free(o);
}

long or_nall(const or_state* o) { return o->nall; }
long or_level_offset(const or_state* o, int n) { return (n >= 1 && n <= o->ng) ? o->off[n] : -1; }

/* ------------------------------------------------------------------ */
/* mg_inject (src/mg_inject.F90:2-64): scalar field fine -> coarse  */
/* levels, all nodes, injection at ii = 2^icrs*i - icrs.               */
/* ------------------------------------------------------------------ */
static void mg_inject(const or_state* o, double* dum, int stride, int comp) {
    int n, i, j, k;
    for (n = 2; n <= o->ng; n++) {
        const int ni = o->img[n], nj = o->jmg[n];
        const int fi = o->img[n - 1], fj = o->jmg[n - 1];
        for (k = 1; k <= o->kmg[n]; k++)
            for (j = 1; j <= nj; j++)
                for (i = 1; i <= ni; i++) {
                    int kk = (1 << o->p.kcrs) * k - o->p.kcrs;
                    int jj = (1 << o->p.jcrs) * j - o->p.jcrs;
                    int ii = (1 << o->p.icrs) * i - o->p.icrs;
                    long ll = o->off[n - 1] + (ii - 1) + (long)fi * ((jj - 1) + (long)fj * (kk - 1));
                    long l = o->off[n] + (i - 1) + (long)ni * ((j - 1) + (long)nj * (k - 1));
                    dum[l * stride + comp] = dum[ll * stride + comp];
                }
    }
}

/* ------------------------------------------------------------------ */
/* mg_metrics (src/mg_metrics.F90:56-300)                               */
/* ------------------------------------------------------------------ */
static void mg_metrics(or_state* o) {
    int n, i, j, k;
    for (n = 1; n <= o->ng; n++) {
        LEVEL(n);
        const long ic = 1, ie = ni, iz = (long)ni * nj;
        double dum1, dum;
        double *x = o->x, *y = o->y, *z = o->z;
        /* csi direction */
        dum1 = o->dc[n];
        dum = one_half * o->dc[n];
        for (k = 1; k <= nk; k++)
            for (j = 1; j <= nj; j++)
                for (i = 2; i <= ni - 1; i++) {
                    long l = off + IX(i, j, k);
                    o->xc[l] = dum * (x[l + ic] - x[l - ic]);
                    o->yc[l] = dum * (y[l + ic] - y[l - ic]);
                    o->zc[l] = dum * (z[l + ic] - z[l - ic]);
                }
        for (k = 1; k <= nk; k++)
            for (j = 1; j <= nj; j++) {
                long l = off + IX(1, j, k);
                o->xc[l] = dum1 * (x[l + ic] - x[l]);
                o->yc[l] = dum1 * (y[l + ic] - y[l]);
                o->zc[l] = dum1 * (z[l + ic] - z[l]);
                l = off + IX(ni, j, k);
                o->xc[l] = dum1 * (x[l] - x[l - ic]);
                o->yc[l] = dum1 * (y[l] - y[l - ic]);
                o->zc[l] = dum1 * (z[l] - z[l - ic]);
            }
        /* eta direction */
        dum1 = o->de[n];
        dum = one_half * o->de[n];
        for (k = 1; k <= nk; k++)
            for (j = 2; j <= nj - 1; j++)
                for (i = 1; i <= ni; i++) {
                    long l = off + IX(i, j, k);
                    o->xe[l] = dum * (x[l + ie] - x[l - ie]);
                    o->ye[l] = dum * (y[l + ie] - y[l - ie]);
                    o->ze[l] = dum * (z[l + ie] - z[l - ie]);
                }
        for (k = 1; k <= nk; k++)
            for (i = 1; i <= ni; i++) {
                long l = off + IX(i, 1, k);
                o->xe[l] = dum1 * (x[l + ie] - x[l]);
                o->ye[l] = dum1 * (y[l + ie] - y[l]);
                o->ze[l] = dum1 * (z[l + ie] - z[l]);
                l = off + IX(i, nj, k);
                o->xe[l] = dum1 * (x[l] - x[l - ie]);
                o->ye[l] = dum1 * (y[l] - y[l - ie]);
                o->ze[l] = dum1 * (z[l] - z[l - ie]);
            }
        /* zet direction */
        dum1 = o->dz[n];
        dum = one_half * o->dz[n];
        for (k = 2; k <= nk - 1; k++)
            for (j = 1; j <= nj; j++)
                for (i = 1; i <= ni; i++) {
                    long l = off + IX(i, j, k);
                    o->xz[l] = dum * (x[l + iz] - x[l - iz]);
                    o->yz[l] = dum * (y[l + iz] - y[l - iz]);
                    o->zz[l] = dum * (z[l + iz] - z[l - iz]);
                }
        for (j = 1; j <= nj; j++)
            for (i = 1; i <= ni; i++) {
                long l = off + IX(i, j, 1);
                o->xz[l] = dum1 * (x[l + iz] - x[l]);
                o->yz[l] = dum1 * (y[l + iz] - y[l]);
                o->zz[l] = dum1 * (z[l + iz] - z[l]);
                l = off + IX(i, j, nk);
                o->xz[l] = dum1 * (x[l] - x[l - iz]);
                o->yz[l] = dum1 * (y[l] - y[l - iz]);
                o->zz[l] = dum1 * (z[l] - z[l - iz]);
            }
        /* Jacobian and inverse transformation (mg_metrics.F90:223-249) */
        for (k = 1; k <= nk; k++)
            for (j = 1; j <= nj; j++)
                for (i = 1; i <= ni; i++) {
                    long l = off + IX(i, j, k);
                    double xc = o->xc[l], xe = o->xe[l], xz = o->xz[l];
                    double yc = o->yc[l], ye = o->ye[l], yz = o->yz[l];
                    double zc = o->zc[l], ze = o->ze[l], zz = o->zz[l];
                    double gj = xc * (ye * zz - yz * ze) + xe * (yz * zc - yc * zz) + xz * (yc * ze - ye * zc);
                    double a = 1.0 / gj;
                    o->aj[l] = a;
                    o->csi[3 * l + 0] = a * (ye * zz - yz * ze);
                    o->csi[3 * l + 1] = a * (xz * ze - xe * zz);
                    o->csi[3 * l + 2] = a * (xe * yz - xz * ye);
                    o->eta[3 * l + 0] = a * (yz * zc - yc * zz);
                    o->eta[3 * l + 1] = a * (xc * zz - xz * zc);
                    o->eta[3 * l + 2] = a * (xz * yc - xc * yz);
                    o->zet[3 * l + 0] = a * (yc * ze - ye * zc);
                    o->zet[3 * l + 1] = a * (xe * zc - xc * ze);
                    o->zet[3 * l + 2] = a * (xc * ye - xe * yc);
                }
    }
    /* dtev on the start level (mg_metrics.F90:262-282) */
    {
        LEVEL(1);
        for (k = 1; k <= nk; k++)
            for (j = 1; j <= nj; j++)
                for (i = 1; i <= ni; i++) {
                    long l = off + IX(i, j, k);
                    double Sc = sqrt(o->xc[l] * o->xc[l] + o->yc[l] * o->yc[l] + o->zc[l] * o->zc[l]);
                    double Se = sqrt(o->xe[l] * o->xe[l] + o->ye[l] * o->ye[l] + o->ze[l] * o->ze[l]);
                    double Sz = sqrt(o->xz[l] * o->xz[l] + o->yz[l] * o->yz[l] + o->zz[l] * o->zz[l]);
                    double SL = dmin2(dmin2(Sc, Se), Sz);
                    double d = SL * dmin2(o->p.cfl2, o->p.vnn2 * o->p.ren * SL);
                    if (o->p.unsteady) d = dmin2(d, o->p.delti * 2.0 / 3.0);
                    o->dtev[l] = d;
                }
    }
}

/* init_outlet_bc (src/init.F90:2750-2891): on every inflow or
 * characteristic-outlet face b (btype 4 or 5) the ratio of the spacings of the first two node pairs
 * off the face in direction bdir(b), for bcond_fm's pressure extrapolation.
 * The k faces test btype == 4 .and. btype == 5 (:2855, :2874), which never
 * holds: their table keeps its allocation contents, zero here. */
static void init_outlet_bc(or_state* o) {
    LEVEL(1);
    for (int b = 0; b < 6; b++) {
        free(o->rat[b]);
        o->rat[b] = NULL;
        if (o->p.btype[b] != 4 && o->p.btype[b] != 5) continue;
        const int na = b < 2 ? nj : ni, nb = nk;  /* in-plane extents (k faces: ni, nj) */
        const int n2 = b < 4 ? nb : nj;
        o->rat[b] = (double*)calloc((size_t)na * n2, sizeof(double));
        if (b >= 4) continue;
        const double* c = o->p.bdir[b] == 1 ? o->x : o->p.bdir[b] == 2 ? o->y : o->z;
        for (int k = 1; k <= nk; k++)
            for (int a = 1; a <= na; a++) {
                long l0, l1, l2;
                if (b == 0) { l0 = IX(1, a, k); l1 = IX(2, a, k); l2 = IX(3, a, k); }
                else if (b == 1) { l0 = IX(ni, a, k); l1 = IX(ni - 1, a, k); l2 = IX(ni - 2, a, k); }
                else if (b == 2) { l0 = IX(a, 1, k); l1 = IX(a, 2, k); l2 = IX(a, 3, k); }
                else { l0 = IX(a, nj, k); l1 = IX(a, nj - 1, k); l2 = IX(a, nj - 2, k); }
                o->rat[b][(a - 1) + (long)na * (k - 1)] =
                    (b % 2 == 0) ? (c[l1] - c[l0]) / (c[l2] - c[l1]) : (c[l0] - c[l1]) / (c[l1] - c[l2]);
            }
    }
}

int or_set_grid(or_state* o, const double* x, const double* y, const double* z) {
    long n1 = o->nl[1];
    memcpy(o->x, x, n1 * sizeof(double));
    memcpy(o->y, y, n1 * sizeof(double));
    memcpy(o->z, z, n1 * sizeof(double));
    mg_inject(o, o->x, 1, 0);
    mg_inject(o, o->y, 1, 0);
    mg_inject(o, o->z, 1, 0);
    mg_metrics(o);
    init_outlet_bc(o);
    return 0;
}

int or_set_wd(or_state* o, const double* wd) {
    memcpy(o->wd, wd, o->nl[1] * sizeof(double));
    return 0;
}

int or_set_state(or_state* o, const double* q, const double* qn, const double* qnm1, const double* xnut) {
    long n1 = o->nl[1];
    if (q) memcpy(o->q, q, 5 * n1 * sizeof(double));
    if (qn) memcpy(o->qn, qn, 5 * n1 * sizeof(double));
    if (qnm1) memcpy(o->qnm1, qnm1, 5 * n1 * sizeof(double));
    if (xnut) memcpy(o->xnut, xnut, n1 * sizeof(double));
    return 0;
}

int or_get_state(const or_state* o, double* q, double* qn, double* qnm1, double* xnut) {
    long n1 = o->nl[1];
    if (q) memcpy(q, o->q, 5 * n1 * sizeof(double));
    if (qn) memcpy(qn, o->qn, 5 * n1 * sizeof(double));
    if (qnm1) memcpy(qnm1, o->qnm1, 5 * n1 * sizeof(double));
    if (xnut) memcpy(xnut, o->xnut, n1 * sizeof(double));
    return 0;
}

static double* or_field(const or_state* o, const char* name, int* ncomp) {
    struct { const char* nm; double* p; int c; } t[] = {
        {"q", o->q, 5}, {"qn", o->qn, 5}, {"qnm1", o->qnm1, 5}, {"qold_mg", o->qold_mg, 5},
        {"rh", o->rh, 4}, {"pk", o->pk, 4}, {"csi", o->csi, 3}, {"eta", o->eta, 3},
        {"zet", o->zet, 3}, {"aj", o->aj, 1}, {"xnut", o->xnut, 1}, {"wd", o->wd, 1},
        {"dtev", o->dtev, 1}, {"x", o->x, 1}, {"y", o->y, 1}, {"z", o->z, 1}};
    size_t i;
    for (i = 0; i < sizeof(t) / sizeof(t[0]); i++)
        if (strcmp(name, t[i].nm) == 0) {
            *ncomp = t[i].c;
            return t[i].p;
        }
    return NULL;
}

int or_get_array(const or_state* o, const char* name, double* out) {
    int c;
    double* p = or_field(o, name, &c);
    if (!p) return -1;
    memcpy(out, p, (size_t)c * o->nall * sizeof(double));
    return 0;
}

int or_set_array(or_state* o, const char* name, const double* in) {
    int c;
    double* p = or_field(o, name, &c);
    if (!p) return -1;
    memcpy(p, in, (size_t)c * o->nall * sizeof(double));
    return 0;
}

/* ------------------------------------------------------------------ */
/* bcond_fm (src/bcond_fm.F90:2-322): fine-grid BCs on q(1:4).         */
/* Its own index range includes the boundary nodes (bcond_fm.F90:60-76) */
/* ------------------------------------------------------------------ */
static void bcond_fm(const or_state* o, double* qv /* level 1, stride 5 */) {
    LEVEL(1);
    int i, j, k, m, b;
    double pfix;
#define QB(m, i, j, k) qv[5 * IX(i, j, k) + (m)-1]
#define RAT(b, a, c, na) o->rat[b][((a)-1) + (long)(na) * ((c)-1)]
    /* i = 1 */
    b = 1;
    switch (o->p.btype[0]) {
        case 1: case 2: case 3:
            for (k = 1; k <= nk; k++)
                for (j = 1; j <= nj; j++)
                    for (m = 1; m <= 4; m++)
                        QB(m, 1, j, k) = o->sa[b - 1][m - 1] * QB(m, 2, j, k) + o->sb[b - 1][m - 1] * QB(m, 3, j, k);
            break;
        case 4: /* inflow: pressure extrapolated on the grid ratio (:95-102) */
            for (k = 1; k <= nk; k++)
                for (j = 1; j <= nj; j++) {
                    const double r = RAT(0, j, k, nj);
                    QB(1, 1, j, k) = (1.0 + r) * QB(1, 2, j, k) - r * QB(1, 3, j, k);
                }
            break;
        case 7: /* periodic wrap (no counterpart in the reference): the image of the
                 * opposite interior plane */
            for (k = 1; k <= nk; k++)
                for (j = 1; j <= nj; j++)
                    for (m = 1; m <= 4; m++) QB(m, 1, j, k) = QB(m, ni - 1, j, k);
            break;
        default: break; /* 0 interface, 6 periodic (commented out, :105-110) */
    }
    /* i = imax */
    b = 2;
    switch (o->p.btype[1]) {
        case 1: case 2: case 3:
            for (k = 1; k <= nk; k++)
                for (j = 1; j <= nj; j++)
                    for (m = 1; m <= 4; m++)
                        QB(m, ni, j, k) = o->sa[b - 1][m - 1] * QB(m, ni - 1, j, k) + o->sb[b - 1][m - 1] * QB(m, ni - 2, j, k);
            break;
        case 4:
            for (k = 1; k <= nk; k++)
                for (j = 1; j <= nj; j++) {
                    const double r = RAT(1, j, k, nj);
                    QB(1, ni, j, k) = (1.0 + r) * QB(1, ni - 1, j, k) - r * QB(1, ni - 2, j, k);
                }
            break;
        case 8: /* free surface (:160-166) */
            for (k = 1; k <= nk; k++)
                for (j = 1; j <= nj; j++)
                    for (m = 1; m <= 4; m++) QB(m, ni, j, k) = QB(m, ni - 1, j, k);
            break;
        case 7: /* periodic wrap (no counterpart in the reference): the image of the
                 * opposite interior plane */
            for (k = 1; k <= nk; k++)
                for (j = 1; j <= nj; j++)
                    for (m = 1; m <= 4; m++) QB(m, ni, j, k) = QB(m, 2, j, k);
            break;
        default: break;
    }
    /* j = 1 */
    b = 3;
    switch (o->p.btype[2]) {
        case 1: case 2: case 3:
            for (k = 1; k <= nk; k++)
                for (i = 1; i <= ni; i++)
                    for (m = 1; m <= 4; m++)
                        QB(m, i, 1, k) = o->sa[b - 1][m - 1] * QB(m, i, 2, k) + o->sb[b - 1][m - 1] * QB(m, i, 3, k);
            break;
        case 4:
            for (k = 1; k <= nk; k++)
                for (i = 1; i <= ni; i++) {
                    const double r = RAT(2, i, k, ni);
                    QB(1, i, 1, k) = (1.0 + r) * QB(1, i, 2, k) - r * QB(1, i, 3, k);
                }
            break;
        case 6: /* plane copy from j = jmax (:197-202) */
            for (k = 1; k <= nk; k++)
                for (i = 1; i <= ni; i++)
                    for (m = 1; m <= 4; m++) QB(m, i, 1, k) = QB(m, i, nj, k);
            break;
        case 7: /* periodic wrap (no counterpart in the reference): the image of the
                 * opposite interior plane */
            for (k = 1; k <= nk; k++)
                for (i = 1; i <= ni; i++)
                    for (m = 1; m <= 4; m++) QB(m, i, 1, k) = QB(m, i, nj - 1, k);
            break;
        default: break;
    }
    /* j = jmax */
    b = 4;
    switch (o->p.btype[3]) {
        case 1: case 2: case 3:
            for (k = 1; k <= nk; k++)
                for (i = 1; i <= ni; i++)
                    for (m = 1; m <= 4; m++)
                        QB(m, i, nj, k) = o->sa[b - 1][m - 1] * QB(m, i, nj - 1, k) + o->sb[b - 1][m - 1] * QB(m, i, nj - 2, k);
            break;
        case 4:
            for (k = 1; k <= nk; k++)
                for (i = 1; i <= ni; i++) {
                    const double r = RAT(3, i, k, ni);
                    QB(1, i, nj, k) = (1.0 + r) * QB(1, i, nj - 1, k) - r * QB(1, i, nj - 2, k);
                }
            break;
        case 6:
            for (k = 1; k <= nk; k++)
                for (i = 1; i <= ni; i++)
                    for (m = 1; m <= 4; m++) QB(m, i, nj, k) = QB(m, i, 1, k);
            break;
        case 7: /* periodic wrap (no counterpart in the reference): the image of the
                 * opposite interior plane */
            for (k = 1; k <= nk; k++)
                for (i = 1; i <= ni; i++)
                    for (m = 1; m <= 4; m++) QB(m, i, nj, k) = QB(m, i, 2, k);
            break;
        default: break;
    }
    /* k = 1 */
    b = 5;
    switch (o->p.btype[4]) {
        case 1: case 2: case 3:
            for (j = 1; j <= nj; j++)
                for (i = 1; i <= ni; i++)
                    for (m = 1; m <= 4; m++)
                        QB(m, i, j, 1) = o->sa[b - 1][m - 1] * QB(m, i, j, 2) + o->sb[b - 1][m - 1] * QB(m, i, j, 3);
            break;
        case 4:
            for (j = 1; j <= nj; j++)
                for (i = 1; i <= ni; i++) {
                    const double r = RAT(4, i, j, ni);
                    QB(1, i, j, 1) = (1.0 + r) * QB(1, i, j, 2) - r * QB(1, i, j, 3);
                }
            break;
        case 6:
            for (j = 1; j <= nj; j++)
                for (i = 1; i <= ni; i++)
                    for (m = 1; m <= 4; m++) QB(m, i, j, 1) = QB(m, i, j, nk);
            break;
        case 7: /* periodic wrap (no counterpart in the reference): the image of the
                 * opposite interior plane */
            for (j = 1; j <= nj; j++)
                for (i = 1; i <= ni; i++)
                    for (m = 1; m <= 4; m++) QB(m, i, j, 1) = QB(m, i, j, nk - 1);
            break;
        default: break;
    }
    /* k = kmax */
    b = 6;
    switch (o->p.btype[5]) {
        case 1: case 2: case 3:
            for (j = 1; j <= nj; j++)
                for (i = 1; i <= ni; i++)
                    for (m = 1; m <= 4; m++)
                        QB(m, i, j, nk) = o->sa[b - 1][m - 1] * QB(m, i, j, nk - 1) + o->sb[b - 1][m - 1] * QB(m, i, j, nk - 2);
            break;
        case 4:
            for (j = 1; j <= nj; j++)
                for (i = 1; i <= ni; i++) {
                    const double r = RAT(5, i, j, ni);
                    QB(1, i, j, nk) = (1.0 + r) * QB(1, i, j, nk - 1) - r * QB(1, i, j, nk - 2);
                }
            break;
        case 6:
            for (j = 1; j <= nj; j++)
                for (i = 1; i <= ni; i++)
                    for (m = 1; m <= 4; m++) QB(m, i, j, nk) = QB(m, i, j, 1);
            break;
        case 7: /* periodic wrap (no counterpart in the reference): the image of the
                 * opposite interior plane */
            for (j = 1; j <= nj; j++)
                for (i = 1; i <= ni; i++)
                    for (m = 1; m <= 4; m++) QB(m, i, j, nk) = QB(m, i, j, 2);
            break;
        default: break;
    }
    /* fixed pressure reference (:310-322) */
    pfix = QB(1, o->p.ifix, o->p.jfix, o->p.kfix);
    for (k = 1; k <= nk; k++)
        for (j = 1; j <= nj; j++)
            for (i = 1; i <= ni; i++) QB(1, i, j, k) = QB(1, i, j, k) - pfix;
    /* zero u on the hub's disk nodes, i = itur1..itur2 (:328-355); the
     * range test covers the whole level, boundaries included */
    for (int t = 0; t < o->n_tur; t++) {
        const desalm_turbine* T = &o->tur[t];
        for (int c = 0; c < T->n_disk; c++) {
            const int di = T->disk_nodes[3 * c], dj = T->disk_nodes[3 * c + 1], dk = T->disk_nodes[3 * c + 2];
            const double radius = T->disk_dat[(dj - 1) + (long)nj * ((dk - 1) + (long)nk * 0)] * o->length_scale;
            if (di >= 1 && di <= ni && dj >= 1 && dj <= nj && dk >= 1 && dk <= nk && radius <= T->hub_rad)
                for (int it = T->itur1; it <= T->itur2; it++) QB(2, it, dj, dk) = 0.0;
        }
    }
#undef QB
#undef RAT
}

/* ------------------------------------------------------------------ */
/* solver_daf and its internal procedures (src/solver_daf.F90)          */
/* ------------------------------------------------------------------ */
typedef struct {
    const or_state* o;
    int n, ni, nj, nk;
    double dc, de, dz;
    int is, ie, js, je, ks, ke, idend;
    double* q;          /* stride 5 */
    const double *qn, *qnm1; /* stride 5 */
    double *rh, *pk;    /* stride 4 */
    const double *csi, *eta, *zet, *aj, *xnut;
    double *qold_af, *diss, *visc, *ucn, *fv, *dtau, *mai, *n1i, *n2i, *mc, *spr;
} daf_t;

#define Q(m, i, j, k) d->q[5 * IX(i, j, k) + (m)-1]
#define QN(m, i, j, k) d->qn[5 * IX(i, j, k) + (m)-1]
#define QNM1(m, i, j, k) d->qnm1[5 * IX(i, j, k) + (m)-1]
#define RH(m, i, j, k) d->rh[4 * IX(i, j, k) + (m)-1]
#define PK(m, i, j, k) d->pk[4 * IX(i, j, k) + (m)-1]
#define CSI(c, i, j, k) d->csi[3 * IX(i, j, k) + (c)-1]
#define ETA(c, i, j, k) d->eta[3 * IX(i, j, k) + (c)-1]
#define ZET(c, i, j, k) d->zet[3 * IX(i, j, k) + (c)-1]
#define AJ(i, j, k) d->aj[IX(i, j, k)]
#define XNUT(i, j, k) d->xnut[IX(i, j, k)]
#define UCN(c, i, j, k) d->ucn[3 * IX(i, j, k) + (c)-1]
#define FV(c, i, j, k) d->fv[3 * IX(i, j, k) + (c)-1]
#define VISC(c, i, j, k) d->visc[3 * IX(i, j, k) + (c)-1]
#define DISS(m, i, j, k) d->diss[4 * IX(i, j, k) + (m)-1]
#define DTAU(i, j, k) d->dtau[IX(i, j, k)]
#define SPR(c, i, j, k) d->spr[3 * IX(i, j, k) + (c)-1]
#define MAT(a, r, c, i, j, k) d->a[16 * IX(i, j, k) + 4 * ((r)-1) + (c)-1]
#define DLV const int ni = d->ni, nj = d->nj, nk = d->nk

/* rhs_contra_j (src/rhs_contra_j.F90:2-31) */
static void rhs_contra_j(daf_t* d) {
    DLV;
    int i, j, k;
    for (k = 1; k <= nk; k++)
        for (j = 1; j <= nj; j++)
            for (i = 1; i <= ni; i++) {
                double tmp = 1.0 / AJ(i, j, k);
                UCN(1, i, j, k) = tmp * (CSI(1, i, j, k) * Q(2, i, j, k) + CSI(2, i, j, k) * Q(3, i, j, k) + CSI(3, i, j, k) * Q(4, i, j, k));
                UCN(2, i, j, k) = tmp * (ETA(1, i, j, k) * Q(2, i, j, k) + ETA(2, i, j, k) * Q(3, i, j, k) + ETA(3, i, j, k) * Q(4, i, j, k));
                UCN(3, i, j, k) = tmp * (ZET(1, i, j, k) * Q(2, i, j, k) + ZET(2, i, j, k) * Q(3, i, j, k) + ZET(3, i, j, k) * Q(4, i, j, k));
            }
}

/* rhs_daf_dtau (src/rhs_daf_dtau.F90:2-100), the live local time step, or
 * with dtau_formula = DESALM_DTAU_MET the viscous limits of met_daf_dtau
 * (src/met_daf_dtau.F90:53-93): vnn * ren / h^2 / g, no nu_t */
static void rhs_daf_dtau(daf_t* d) {
    DLV;
    int i, j, k;
    const double cfl = d->o->p.cfl1, vnn = d->o->p.vnn1, rei = 1.0 / d->o->p.ren, ren = d->o->p.ren;
    const int met = d->o->p.dtau_formula == DESALM_DTAU_MET;
    const double dcsq = d->dc * d->dc, desq = d->de * d->de, dzsq = d->dz * d->dz;
    for (k = 1; k <= nk; k++)
        for (j = 1; j <= nj; j++)
            for (i = 1; i <= ni; i++) {
                double g1 = CSI(1, i, j, k) * CSI(1, i, j, k) + CSI(2, i, j, k) * CSI(2, i, j, k) + CSI(3, i, j, k) * CSI(3, i, j, k);
                double g2 = ETA(1, i, j, k) * ETA(1, i, j, k) + ETA(2, i, j, k) * ETA(2, i, j, k) + ETA(3, i, j, k) * ETA(3, i, j, k);
                double g3 = ZET(1, i, j, k) * ZET(1, i, j, k) + ZET(2, i, j, k) * ZET(2, i, j, k) + ZET(3, i, j, k) * ZET(3, i, j, k);
                double e1 = sqrt(g1), e2 = sqrt(g2), e3 = sqrt(g3);
                double dti1 = cfl / e1 / d->dc, dti2 = cfl / e2 / d->de, dti3 = cfl / e3 / d->dz;
                double dtv1, dtv2, dtv3;
                if (met) {
                    dtv1 = vnn * ren / dcsq / g1;
                    dtv2 = vnn * ren / desq / g2;
                    dtv3 = vnn * ren / dzsq / g3;
                } else {
                    double ret = rei + XNUT(i, j, k);
                    dtv1 = vnn / ret / dcsq / g1;
                    dtv2 = vnn / ret / desq / g2;
                    dtv3 = vnn / ret / dzsq / g3;
                }
                double t = dmin2(dti1, dti2);
                t = dmin2(t, dti3);
                t = dmin2(t, dtv1);
                t = dmin2(t, dtv2);
                t = dmin2(t, dtv3);
                DTAU(i, j, k) = t;
            }
}

/* rhs_viscous (src/rhs_viscous.F90:2-333) */
static void rhs_viscous(daf_t* d) {
    DLV;
    int i, j, k, c;
    const double rei = 1.0 / d->o->p.ren;
    const double dc = d->dc, de = d->de, dz = d->dz;
    const double dc4 = one_fourth * dc, de4 = one_fourth * de, dz4 = one_fourth * dz;
    long N = (long)ni * nj * nk;
    double* ret = zalloc(N);
    memset(d->fv, 0, 3 * N * sizeof(double));
    memset(d->visc, 0, 3 * N * sizeof(double));
    for (k = 1; k <= nk; k++)
        for (j = 1; j <= nj; j++)
            for (i = 1; i <= ni; i++) ret[IX(i, j, k)] = rei + XNUT(i, j, k);
#define RET(i, j, k) ret[IX(i, j, k)]
    /* csi faces */
    for (k = d->ks; k <= d->ke; k++)
        for (j = d->js; j <= d->je; j++)
            for (i = d->is - 1; i <= d->ie; i++) {
                double ucs = dc * (Q(2, i + 1, j, k) - Q(2, i, j, k));
                double vcs = dc * (Q(3, i + 1, j, k) - Q(3, i, j, k));
                double wcs = dc * (Q(4, i + 1, j, k) - Q(4, i, j, k));
                double uet = de4 * (Q(2, i + 1, j + 1, k) - Q(2, i + 1, j - 1, k) + Q(2, i, j + 1, k) - Q(2, i, j - 1, k));
                double vet = de4 * (Q(3, i + 1, j + 1, k) - Q(3, i + 1, j - 1, k) + Q(3, i, j + 1, k) - Q(3, i, j - 1, k));
                double wet = de4 * (Q(4, i + 1, j + 1, k) - Q(4, i + 1, j - 1, k) + Q(4, i, j + 1, k) - Q(4, i, j - 1, k));
                double uze = dz4 * (Q(2, i + 1, j, k + 1) - Q(2, i + 1, j, k - 1) + Q(2, i, j, k + 1) - Q(2, i, j, k - 1));
                double vze = dz4 * (Q(3, i + 1, j, k + 1) - Q(3, i + 1, j, k - 1) + Q(3, i, j, k + 1) - Q(3, i, j, k - 1));
                double wze = dz4 * (Q(4, i + 1, j, k + 1) - Q(4, i + 1, j, k - 1) + Q(4, i, j, k + 1) - Q(4, i, j, k - 1));
                double a1 = CSI(1, i + 1, j, k) + CSI(1, i, j, k), a2 = CSI(2, i + 1, j, k) + CSI(2, i, j, k), a3 = CSI(3, i + 1, j, k) + CSI(3, i, j, k);
                double b1 = ETA(1, i + 1, j, k) + ETA(1, i, j, k), b2 = ETA(2, i + 1, j, k) + ETA(2, i, j, k), b3 = ETA(3, i + 1, j, k) + ETA(3, i, j, k);
                double c1 = ZET(1, i + 1, j, k) + ZET(1, i, j, k), c2 = ZET(2, i + 1, j, k) + ZET(2, i, j, k), c3 = ZET(3, i + 1, j, k) + ZET(3, i, j, k);
                double uu1 = a1 * ucs + b1 * uet + c1 * uze, vv1 = a1 * vcs + b1 * vet + c1 * vze, ww1 = a1 * wcs + b1 * wet + c1 * wze;
                double uu2 = a2 * ucs + b2 * uet + c2 * uze, vv2 = a2 * vcs + b2 * vet + c2 * vze, ww2 = a2 * wcs + b2 * wet + c2 * wze;
                double uu3 = a3 * ucs + b3 * uet + c3 * uze, vv3 = a3 * vcs + b3 * vet + c3 * vze, ww3 = a3 * wcs + b3 * wet + c3 * wze;
                double c11 = a1 * a1 + a2 * a2 + a3 * a3, c12 = a1 * b1 + a2 * b2 + a3 * b3, c13 = a1 * c1 + a2 * c2 + a3 * c3;
                double vtj = one_fourth * (RET(i + 1, j, k) + RET(i, j, k)) / (AJ(i + 1, j, k) + AJ(i, j, k));
                FV(1, i, j, k) = vtj * (a1 * uu1 + a2 * vv1 + a3 * ww1 + c11 * ucs + c12 * uet + c13 * uze);
                FV(2, i, j, k) = vtj * (a1 * uu2 + a2 * vv2 + a3 * ww2 + c11 * vcs + c12 * vet + c13 * vze);
                FV(3, i, j, k) = vtj * (a1 * uu3 + a2 * vv3 + a3 * ww3 + c11 * wcs + c12 * wet + c13 * wze);
            }
    for (k = d->ks; k <= d->ke; k++)
        for (j = d->js; j <= d->je; j++)
            for (i = d->is; i <= d->ie; i++)
                for (c = 1; c <= 3; c++) VISC(c, i, j, k) = dc * (FV(c, i, j, k) - FV(c, i - 1, j, k));
    /* eta faces */
    for (k = d->ks; k <= d->ke; k++)
        for (j = d->js - 1; j <= d->je; j++)
            for (i = d->is; i <= d->ie; i++) {
                double ucs = dc4 * (Q(2, i + 1, j + 1, k) - Q(2, i - 1, j + 1, k) + Q(2, i + 1, j, k) - Q(2, i - 1, j, k));
                double vcs = dc4 * (Q(3, i + 1, j + 1, k) - Q(3, i - 1, j + 1, k) + Q(3, i + 1, j, k) - Q(3, i - 1, j, k));
                double wcs = dc4 * (Q(4, i + 1, j + 1, k) - Q(4, i - 1, j + 1, k) + Q(4, i + 1, j, k) - Q(4, i - 1, j, k));
                double uet = de * (Q(2, i, j + 1, k) - Q(2, i, j, k));
                double vet = de * (Q(3, i, j + 1, k) - Q(3, i, j, k));
                double wet = de * (Q(4, i, j + 1, k) - Q(4, i, j, k));
                double uze = dz4 * (Q(2, i, j + 1, k + 1) - Q(2, i, j + 1, k - 1) + Q(2, i, j, k + 1) - Q(2, i, j, k - 1));
                double vze = dz4 * (Q(3, i, j + 1, k + 1) - Q(3, i, j + 1, k - 1) + Q(3, i, j, k + 1) - Q(3, i, j, k - 1));
                double wze = dz4 * (Q(4, i, j + 1, k + 1) - Q(4, i, j + 1, k - 1) + Q(4, i, j, k + 1) - Q(4, i, j, k - 1));
                double a1 = CSI(1, i, j + 1, k) + CSI(1, i, j, k), a2 = CSI(2, i, j + 1, k) + CSI(2, i, j, k), a3 = CSI(3, i, j + 1, k) + CSI(3, i, j, k);
                double b1 = ETA(1, i, j + 1, k) + ETA(1, i, j, k), b2 = ETA(2, i, j + 1, k) + ETA(2, i, j, k), b3 = ETA(3, i, j + 1, k) + ETA(3, i, j, k);
                double c1 = ZET(1, i, j + 1, k) + ZET(1, i, j, k), c2 = ZET(2, i, j + 1, k) + ZET(2, i, j, k), c3 = ZET(3, i, j + 1, k) + ZET(3, i, j, k);
                double uu1 = a1 * ucs + b1 * uet + c1 * uze, vv1 = a1 * vcs + b1 * vet + c1 * vze, ww1 = a1 * wcs + b1 * wet + c1 * wze;
                double uu2 = a2 * ucs + b2 * uet + c2 * uze, vv2 = a2 * vcs + b2 * vet + c2 * vze, ww2 = a2 * wcs + b2 * wet + c2 * wze;
                double uu3 = a3 * ucs + b3 * uet + c3 * uze, vv3 = a3 * vcs + b3 * vet + c3 * vze, ww3 = a3 * wcs + b3 * wet + c3 * wze;
                double c21 = b1 * a1 + b2 * a2 + b3 * a3, c22 = b1 * b1 + b2 * b2 + b3 * b3, c23 = b1 * c1 + b2 * c2 + b3 * c3;
                double vtj = one_fourth * (RET(i, j + 1, k) + RET(i, j, k)) / (AJ(i, j + 1, k) + AJ(i, j, k));
                FV(1, i, j, k) = vtj * (b1 * uu1 + b2 * vv1 + b3 * ww1 + c21 * ucs + c22 * uet + c23 * uze);
                FV(2, i, j, k) = vtj * (b1 * uu2 + b2 * vv2 + b3 * ww2 + c21 * vcs + c22 * vet + c23 * vze);
                FV(3, i, j, k) = vtj * (b1 * uu3 + b2 * vv3 + b3 * ww3 + c21 * wcs + c22 * wet + c23 * wze);
            }
    for (k = d->ks; k <= d->ke; k++)
        for (j = d->js; j <= d->je; j++)
            for (i = d->is; i <= d->ie; i++)
                for (c = 1; c <= 3; c++) VISC(c, i, j, k) = VISC(c, i, j, k) + de * (FV(c, i, j, k) - FV(c, i, j - 1, k));
    /* zet faces */
    for (k = d->ks - 1; k <= d->ke; k++)
        for (j = d->js; j <= d->je; j++)
            for (i = d->is; i <= d->ie; i++) {
                double ucs = dc4 * (Q(2, i + 1, j, k + 1) - Q(2, i - 1, j, k + 1) + Q(2, i + 1, j, k) - Q(2, i - 1, j, k));
                double vcs = dc4 * (Q(3, i + 1, j, k + 1) - Q(3, i - 1, j, k + 1) + Q(3, i + 1, j, k) - Q(3, i - 1, j, k));
                double wcs = dc4 * (Q(4, i + 1, j, k + 1) - Q(4, i - 1, j, k + 1) + Q(4, i + 1, j, k) - Q(4, i - 1, j, k));
                double uet = de4 * (Q(2, i, j + 1, k + 1) - Q(2, i, j - 1, k + 1) + Q(2, i, j + 1, k) - Q(2, i, j - 1, k));
                double vet = de4 * (Q(3, i, j + 1, k + 1) - Q(3, i, j - 1, k + 1) + Q(3, i, j + 1, k) - Q(3, i, j - 1, k));
                double wet = de4 * (Q(4, i, j + 1, k + 1) - Q(4, i, j - 1, k + 1) + Q(4, i, j + 1, k) - Q(4, i, j - 1, k));
                double uze = dz * (Q(2, i, j, k + 1) - Q(2, i, j, k));
                double vze = dz * (Q(3, i, j, k + 1) - Q(3, i, j, k));
                double wze = dz * (Q(4, i, j, k + 1) - Q(4, i, j, k));
                double a1 = CSI(1, i, j, k + 1) + CSI(1, i, j, k), a2 = CSI(2, i, j, k + 1) + CSI(2, i, j, k), a3 = CSI(3, i, j, k + 1) + CSI(3, i, j, k);
                double b1 = ETA(1, i, j, k + 1) + ETA(1, i, j, k), b2 = ETA(2, i, j, k + 1) + ETA(2, i, j, k), b3 = ETA(3, i, j, k + 1) + ETA(3, i, j, k);
                double c1 = ZET(1, i, j, k + 1) + ZET(1, i, j, k), c2 = ZET(2, i, j, k + 1) + ZET(2, i, j, k), c3 = ZET(3, i, j, k + 1) + ZET(3, i, j, k);
                double uu1 = a1 * ucs + b1 * uet + c1 * uze, vv1 = a1 * vcs + b1 * vet + c1 * vze, ww1 = a1 * wcs + b1 * wet + c1 * wze;
                double uu2 = a2 * ucs + b2 * uet + c2 * uze, vv2 = a2 * vcs + b2 * vet + c2 * vze, ww2 = a2 * wcs + b2 * wet + c2 * wze;
                double uu3 = a3 * ucs + b3 * uet + c3 * uze, vv3 = a3 * vcs + b3 * vet + c3 * vze, ww3 = a3 * wcs + b3 * wet + c3 * wze;
                double c31 = c1 * a1 + c2 * a2 + c3 * a3, c32 = c1 * b1 + c2 * b2 + c3 * b3, c33 = c1 * c1 + c2 * c2 + c3 * c3;
                double vtj = one_fourth * (RET(i, j, k + 1) + RET(i, j, k)) / (AJ(i, j, k + 1) + AJ(i, j, k));
                FV(1, i, j, k) = vtj * (c1 * uu1 + c2 * vv1 + c3 * ww1 + c31 * ucs + c32 * uet + c33 * uze);
                FV(2, i, j, k) = vtj * (c1 * uu2 + c2 * vv2 + c3 * ww2 + c31 * vcs + c32 * vet + c33 * vze);
                FV(3, i, j, k) = vtj * (c1 * uu3 + c2 * vv3 + c3 * ww3 + c31 * wcs + c32 * wet + c33 * wze);
            }
    for (k = d->ks; k <= d->ke; k++)
        for (j = d->js; j <= d->je; j++)
            for (i = d->is; i <= d->ie; i++)
                for (c = 1; c <= 3; c++) VISC(c, i, j, k) = VISC(c, i, j, k) + dz * (FV(c, i, j, k) - FV(c, i, j, k - 1));
#undef RET
    free(ret);
}

/* rhs_diss_p (src/rhs_diss_p.F90:2-205) */
static void rhs_diss_p(daf_t* d) {
    DLV;
    int i, j, k;
    const double dcsq = d->dc * d->dc, desq = d->de * d->de, dzsq = d->dz * d->dz;
    long N = (long)ni * nj * nk;
    memset(d->fv, 0, 3 * N * sizeof(double));
    memset(d->diss, 0, 4 * N * sizeof(double));
    for (k = d->ks - 1; k <= d->ke + 1; k++)
        for (j = d->js - 1; j <= d->je + 1; j++)
            for (i = d->is; i <= d->ie; i++) {
                double g1 = CSI(1, i, j, k) * CSI(1, i, j, k) + CSI(2, i, j, k) * CSI(2, i, j, k) + CSI(3, i, j, k) * CSI(3, i, j, k);
                double tmp1 = g1 * DTAU(i, j, k) / AJ(i, j, k);
                double tmp2 = dcsq * (Q(1, i + 1, j, k) - 2.0 * Q(1, i, j, k) + Q(1, i - 1, j, k));
                FV(1, i, j, k) = tmp1 * tmp2;
            }
    for (k = d->ks - 1; k <= d->ke + 1; k++)
        for (j = d->js; j <= d->je; j++)
            for (i = d->is - 1; i <= d->ie + 1; i++) {
                double g2 = ETA(1, i, j, k) * ETA(1, i, j, k) + ETA(2, i, j, k) * ETA(2, i, j, k) + ETA(3, i, j, k) * ETA(3, i, j, k);
                double tmp1 = g2 * DTAU(i, j, k) / AJ(i, j, k);
                double tmp2 = desq * (Q(1, i, j + 1, k) - 2.0 * Q(1, i, j, k) + Q(1, i, j - 1, k));
                FV(2, i, j, k) = tmp1 * tmp2;
            }
    for (k = d->ks; k <= d->ke; k++)
        for (j = d->js - 1; j <= d->je + 1; j++)
            for (i = d->is - 1; i <= d->ie + 1; i++) {
                double g3 = ZET(1, i, j, k) * ZET(1, i, j, k) + ZET(2, i, j, k) * ZET(2, i, j, k) + ZET(3, i, j, k) * ZET(3, i, j, k);
                double tmp1 = g3 * DTAU(i, j, k) / AJ(i, j, k);
                double tmp2 = dzsq * (Q(1, i, j, k + 1) - 2.0 * Q(1, i, j, k) + Q(1, i, j, k - 1));
                FV(3, i, j, k) = tmp1 * tmp2;
            }
    /* one-sided copies at the domain boundaries (:124-129) */
    for (k = 1; k <= nk; k++)
        for (j = 1; j <= nj; j++) {
            FV(1, d->is - 1, j, k) = FV(1, d->is, j, k);
            FV(1, d->ie + 1, j, k) = FV(1, d->ie, j, k);
        }
    for (k = 1; k <= nk; k++)
        for (i = 1; i <= ni; i++) {
            FV(2, i, d->js - 1, k) = FV(2, i, d->js, k);
            FV(2, i, d->je + 1, k) = FV(2, i, d->je, k);
        }
    for (j = 1; j <= nj; j++)
        for (i = 1; i <= ni; i++) {
            FV(3, i, j, d->ks - 1) = FV(3, i, j, d->ks);
            FV(3, i, j, d->ke + 1) = FV(3, i, j, d->ke);
        }
    for (k = d->ks; k <= d->ke; k++)
        for (j = d->js; j <= d->je; j++)
            for (i = d->is; i <= d->ie; i++)
                DISS(1, i, j, k) = d->o->p.pdiss_coef *
                                   ((FV(1, i + 1, j, k) - 2.0 * FV(1, i, j, k) + FV(1, i - 1, j, k)) +
                                    (FV(2, i, j + 1, k) - 2.0 * FV(2, i, j, k) + FV(2, i, j - 1, k)) +
                                    (FV(3, i, j, k + 1) - 2.0 * FV(3, i, j, k) + FV(3, i, j, k - 1)));
}

/* rhs_flux_sans_convec (src/rhs_flux_sans_convec.F90:2-140) */
static void rhs_flux_sans_convec(daf_t* d) {
    DLV;
    int i, j, k, c;
    const double dc2 = one_half * d->dc, de2 = one_half * d->de, dz2 = one_half * d->dz;
    long N = (long)ni * nj * nk;
    double* p = zalloc(N);
#define P(i, j, k) p[IX(i, j, k)]
    for (k = 1; k <= nk; k++)
        for (j = 1; j <= nj; j++)
            for (i = 1; i <= ni; i++) P(i, j, k) = Q(1, i, j, k) / AJ(i, j, k);
    for (k = d->ks; k <= d->ke; k++)
        for (j = d->js; j <= d->je; j++)
            for (i = d->is; i <= d->ie; i++)
                for (c = 1; c <= 3; c++)
                    RH(c + 1, i, j, k) = dc2 * (P(i + 1, j, k) * CSI(c, i + 1, j, k) - P(i - 1, j, k) * CSI(c, i - 1, j, k));
    for (k = d->ks; k <= d->ke; k++)
        for (j = d->js; j <= d->je; j++)
            for (i = d->is; i <= d->ie; i++)
                for (c = 1; c <= 3; c++)
                    RH(c + 1, i, j, k) = RH(c + 1, i, j, k) + de2 * (P(i, j + 1, k) * ETA(c, i, j + 1, k) - P(i, j - 1, k) * ETA(c, i, j - 1, k));
    for (k = d->ks; k <= d->ke; k++)
        for (j = d->js; j <= d->je; j++)
            for (i = d->is; i <= d->ie; i++)
                for (c = 1; c <= 3; c++)
                    RH(c + 1, i, j, k) = RH(c + 1, i, j, k) + dz2 * (P(i, j, k + 1) * ZET(c, i, j, k + 1) - P(i, j, k - 1) * ZET(c, i, j, k - 1));
    for (k = d->ks; k <= d->ke; k++)
        for (j = d->js; j <= d->je; j++)
            for (i = d->is; i <= d->ie; i++) {
                double dudx = dc2 * (UCN(1, i + 1, j, k) - UCN(1, i - 1, j, k));
                double dvdy = de2 * (UCN(2, i, j + 1, k) - UCN(2, i, j - 1, k));
                double dwdz = dz2 * (UCN(3, i, j, k + 1) - UCN(3, i, j, k - 1));
                RH(1, i, j, k) = dudx + dvdy + dwdz;
            }
#undef P
    free(p);
}

/* rhs_convec_quick (src/rhs_convec_quick.F90:2-438) */
static void rhs_convec_quick(daf_t* d, int decide_upwind) {
    DLV;
    int i, j, k, m, dir;
    const double coef = 0.125;
    long N = (long)ni * nj * nk;
    double* up = zalloc(N);
    double* um = zalloc(N);
    const int ista = d->is + 1, iend = d->ie - 1, jsta = d->js + 1, jend = d->je - 1, ksta = d->ks + 1, kend = d->ke - 1;
#define UP(i, j, k) up[IX(i, j, k)]
#define UM(i, j, k) um[IX(i, j, k)]
    for (dir = 1; dir <= 3; dir++) {
        const int di = dir == 1, dj = dir == 2, dk = dir == 3;
        const double h = dir == 1 ? d->dc : dir == 2 ? d->de : d->dz;
        for (k = 1; k <= nk; k++)
            for (j = 1; j <= nj; j++)
                for (i = 1; i <= ni; i++) {
                    double ucon = one_half * UCN(dir, i, j, k);
                    UP(i, j, k) = ucon + fabs(ucon);
                    UM(i, j, k) = ucon - fabs(ucon);
                }
        if (decide_upwind == 1) {
            for (k = d->ks; k <= d->ke; k++)
                for (j = d->js; j <= d->je; j++)
                    for (i = d->is; i <= d->ie; i++)
                        for (m = 2; m <= 4; m++)
                            FV(m - 1, i, j, k) = h * (UP(i, j, k) * (Q(m, i, j, k) - Q(m, i - di, j - dj, k - dk)) +
                                                      UM(i, j, k) * (Q(m, i + di, j + dj, k + dk) - Q(m, i, j, k)));
        } else {
            int i0 = d->is, i1 = d->ie, j0 = d->js, j1 = d->je, k0 = d->ks, k1 = d->ke;
            if (dir == 1) { i0 = ista; i1 = iend; }
            if (dir == 2) { j0 = jsta; j1 = jend; }
            if (dir == 3) { k0 = ksta; k1 = kend; }
            for (k = k0; k <= k1; k++)
                for (j = j0; j <= j1; j++)
                    for (i = i0; i <= i1; i++)
                        for (m = 2; m <= 4; m++)
                            FV(m - 1, i, j, k) =
                                coef * h *
                                (UP(i, j, k) * (3.0 * Q(m, i + di, j + dj, k + dk) + 3.0 * Q(m, i, j, k) -
                                                 7.0 * Q(m, i - di, j - dj, k - dk) + Q(m, i - 2 * di, j - 2 * dj, k - 2 * dk)) +
                                 UM(i, j, k) * (-3.0 * Q(m, i - di, j - dj, k - dk) - 3.0 * Q(m, i, j, k) +
                                                 7.0 * Q(m, i + di, j + dj, k + dk) - Q(m, i + 2 * di, j + 2 * dj, k + 2 * dk)));
            /* first-order upwind at the first and last interior nodes (:140-161) */
            if (dir == 1) {
                for (k = d->ks; k <= d->ke; k++)
                    for (j = d->js; j <= d->je; j++) {
                        int ib, ii;
                        for (ib = 0; ib < 2; ib++) {
                            ii = ib == 0 ? d->is : d->ie;
                            for (m = 2; m <= 4; m++)
                                FV(m - 1, ii, j, k) = h * (UP(ii, j, k) * (Q(m, ii, j, k) - Q(m, ii - 1, j, k)) +
                                                           UM(ii, j, k) * (Q(m, ii + 1, j, k) - Q(m, ii, j, k)));
                        }
                    }
            } else if (dir == 2) {
                for (k = d->ks; k <= d->ke; k++)
                    for (i = d->is; i <= d->ie; i++) {
                        int jb, jj;
                        for (jb = 0; jb < 2; jb++) {
                            jj = jb == 0 ? d->js : d->je;
                            for (m = 2; m <= 4; m++)
                                FV(m - 1, i, jj, k) = h * (UP(i, jj, k) * (Q(m, i, jj, k) - Q(m, i, jj - 1, k)) +
                                                           UM(i, jj, k) * (Q(m, i, jj + 1, k) - Q(m, i, jj, k)));
                        }
                    }
            } else {
                for (j = d->js; j <= d->je; j++)
                    for (i = d->is; i <= d->ie; i++) {
                        int kb, kk;
                        for (kb = 0; kb < 2; kb++) {
                            kk = kb == 0 ? d->ks : d->ke;
                            for (m = 2; m <= 4; m++)
                                FV(m - 1, i, j, kk) = h * (UP(i, j, kk) * (Q(m, i, j, kk) - Q(m, i, j, kk - 1)) +
                                                           UM(i, j, kk) * (Q(m, i, j, kk + 1) - Q(m, i, j, kk)));
                        }
                    }
            }
        }
        for (k = d->ks; k <= d->ke; k++)
            for (j = d->js; j <= d->je; j++)
                for (i = d->is; i <= d->ie; i++)
                    for (m = 2; m <= 4; m++) RH(m, i, j, k) = RH(m, i, j, k) + FV(m - 1, i, j, k);
    }
    /* diss(2:4) = 0 (:433) */
    for (k = 1; k <= nk; k++)
        for (j = 1; j <= nj; j++)
            for (i = 1; i <= ni; i++)
                for (m = 2; m <= 4; m++) DISS(m, i, j, k) = 0.0;
#undef UP
#undef UM
    free(up);
    free(um);
}

/* rhs_unst_visc_diss (src/rhs_unst_visc_diss.F90:2-52) */
static void rhs_unst_visc_diss(daf_t* d) {
    DLV;
    int i, j, k, m;
    const double es = d->o->e_source, delti = d->o->p.delti;
    const double* fb = d->o->p.body_force;
    (void)nk;
    for (k = d->ks; k <= d->ke; k++)
        for (j = d->js; j <= d->je; j++)
            for (i = d->is; i <= d->ie; i++) {
                double dtaj = 2.0 * delti * AJ(i, j, k);
                double dqdt[3];
                for (m = 2; m <= 4; m++)
                    dqdt[m - 2] = es * (3.0 * Q(m, i, j, k) - 4.0 * QN(m, i, j, k) + QNM1(m, i, j, k)) / dtaj;
                RH(1, i, j, k) = RH(1, i, j, k) + DISS(1, i, j, k);
                for (m = 2; m <= 4; m++)
                    RH(m, i, j, k) = RH(m, i, j, k) + DISS(m, i, j, k) + dqdt[m - 2] - VISC(m - 1, i, j, k);
                /* the body force (no counterpart in the reference): -f / J */
                if (fb[0] != 0.0 || fb[1] != 0.0 || fb[2] != 0.0)
                    for (m = 2; m <= 4; m++) RH(m, i, j, k) = RH(m, i, j, k) - fb[m - 2] / AJ(i, j, k);
            }
}

/* rhs_pk (src/rhs_pk.F90:2-69): FAS forcing on interior +/- 1 */
static void rhs_pk(daf_t* d, int calc) {
    DLV;
    int i, j, k, m;
    (void)nk;
    if (calc == 1)
        for (k = d->ks - 1; k <= d->ke + 1; k++)
            for (j = d->js - 1; j <= d->je + 1; j++)
                for (i = d->is - 1; i <= d->ie + 1; i++)
                    for (m = 1; m <= 4; m++) PK(m, i, j, k) = PK(m, i, j, k) - RH(m, i, j, k);
    for (k = d->ks - 1; k <= d->ke + 1; k++)
        for (j = d->js - 1; j <= d->je + 1; j++)
            for (i = d->is - 1; i <= d->ie + 1; i++)
                for (m = 1; m <= 4; m++) RH(m, i, j, k) = RH(m, i, j, k) + PK(m, i, j, k);
}

static void zero_rh_boundaries(daf_t* d) {
    DLV;
    int i, j, k, m;
    for (k = 1; k <= nk; k++)
        for (j = 1; j <= nj; j++)
            for (m = 1; m <= 4; m++) RH(m, d->is - 1, j, k) = 0.0, RH(m, d->ie + 1, j, k) = 0.0;
    for (k = 1; k <= nk; k++)
        for (i = 1; i <= ni; i++)
            for (m = 1; m <= 4; m++) RH(m, i, d->js - 1, k) = 0.0, RH(m, i, d->je + 1, k) = 0.0;
    for (j = 1; j <= nj; j++)
        for (i = 1; i <= ni; i++)
            for (m = 1; m <= 4; m++) RH(m, i, j, d->ks - 1) = 0.0, RH(m, i, j, d->ke + 1) = 0.0;
}

/* ------------------------------------------------------------------ */
/* Characteristic outlet (btype 5) of an i face on the fine level:      */
/* mg_brhs (src/mg_brhs.F90:2-108 with brhs_viscous, brhs_diss_p,        */
/* brhs_flux_sans_convec, brhs_convec_quick(2), brhs_unst_visc_diss)     */
/* builds rh on the boundary plane from its eta and zeta derivatives    */
/* only; nonreflect_ibc (src/nonreflect_ibc.F90:1-244) then adds the     */
/* i-direction terms as characteristic amplitudes, the incoming ones    */
/* set to zero.  ip: the plane (1 or ni); inout: 1 or 2.                */
/* ------------------------------------------------------------------ */
static void mg_brhs_plane(daf_t* d, int ip) {
    DLV;
    const or_state* o = d->o;
    const double de = d->de, dz = d->dz;
    const int js = 2, je = nj - 1, ks = 2, ke = nk - 1;
    const long np = (long)nj * nk;
    double* visc = zalloc(3 * np);
    double* diss = zalloc(4 * np);
    double* fv = zalloc(3 * np);
    double* ret = zalloc(np);
    double* aux = zalloc(2 * np);
    double* pp = zalloc(np);
    double* up = zalloc(np);
    double* um = zalloc(np);
    double* rd = zalloc(3 * np);
#define P2(j, k) ((long)((j)-1) + (long)nj * ((k)-1))
#define PQ(m, j, k) Q(m, ip, j, k)
    int j, k, c;
    /* brhs_viscous */
    const double rei = 1.0 / o->p.ren;
    for (k = 1; k <= nk; k++)
        for (j = 1; j <= nj; j++) ret[P2(j, k)] = rei + XNUT(ip, j, k);
    const double de4 = pt25 * de, dz4 = pt25 * dz;
    for (k = ks; k <= ke; k++)
        for (j = js - 1; j <= je; j++) {
            const double uet = de * (PQ(2, j + 1, k) - PQ(2, j, k));
            const double vet = de * (PQ(3, j + 1, k) - PQ(3, j, k));
            const double wet = de * (PQ(4, j + 1, k) - PQ(4, j, k));
            const double uze = dz4 * (PQ(2, j + 1, k + 1) - PQ(2, j + 1, k - 1) + PQ(2, j, k + 1) - PQ(2, j, k - 1));
            const double vze = dz4 * (PQ(3, j + 1, k + 1) - PQ(3, j + 1, k - 1) + PQ(3, j, k + 1) - PQ(3, j, k - 1));
            const double wze = dz4 * (PQ(4, j + 1, k + 1) - PQ(4, j + 1, k - 1) + PQ(4, j, k + 1) - PQ(4, j, k - 1));
            const double b1 = ETA(1, ip, j + 1, k) + ETA(1, ip, j, k);
            const double b2 = ETA(2, ip, j + 1, k) + ETA(2, ip, j, k);
            const double b3 = ETA(3, ip, j + 1, k) + ETA(3, ip, j, k);
            const double c1 = ZET(1, ip, j + 1, k) + ZET(1, ip, j, k);
            const double c2 = ZET(2, ip, j + 1, k) + ZET(2, ip, j, k);
            const double c3 = ZET(3, ip, j + 1, k) + ZET(3, ip, j, k);
            const double uu1 = b1 * uet + c1 * uze, vv1 = b1 * vet + c1 * vze, ww1 = b1 * wet + c1 * wze;
            const double uu2 = b2 * uet + c2 * uze, vv2 = b2 * vet + c2 * vze, ww2 = b2 * wet + c2 * wze;
            const double uu3 = b3 * uet + c3 * uze, vv3 = b3 * vet + c3 * vze, ww3 = b3 * wet + c3 * wze;
            const double c22 = b1 * b1 + b2 * b2 + b3 * b3;
            const double c23 = b1 * c1 + b2 * c2 + b3 * c3;
            const double vtj = one_fourth * (ret[P2(j + 1, k)] + ret[P2(j, k)]) / (AJ(ip, j + 1, k) + AJ(ip, j, k));
            fv[3 * P2(j, k) + 0] = vtj * (b1 * uu1 + b2 * vv1 + b3 * ww1 + c22 * uet + c23 * uze);
            fv[3 * P2(j, k) + 1] = vtj * (b1 * uu2 + b2 * vv2 + b3 * ww2 + c22 * vet + c23 * vze);
            fv[3 * P2(j, k) + 2] = vtj * (b1 * uu3 + b2 * vv3 + b3 * ww3 + c22 * wet + c23 * wze);
        }
    for (k = ks; k <= ke; k++)
        for (j = js; j <= je; j++)
            for (c = 0; c < 3; c++) visc[3 * P2(j, k) + c] = de * (fv[3 * P2(j, k) + c] - fv[3 * P2(j - 1, k) + c]);
    for (k = ks - 1; k <= ke; k++)
        for (j = js; j <= je; j++) {
            const double uet = de4 * (PQ(2, j + 1, k + 1) - PQ(2, j - 1, k + 1) + PQ(2, j + 1, k) - PQ(2, j - 1, k));
            const double vet = de4 * (PQ(3, j + 1, k + 1) - PQ(3, j - 1, k + 1) + PQ(3, j + 1, k) - PQ(3, j - 1, k));
            const double wet = de4 * (PQ(4, j + 1, k + 1) - PQ(4, j - 1, k + 1) + PQ(4, j + 1, k) - PQ(4, j - 1, k));
            const double uze = dz * (PQ(2, j, k + 1) - PQ(2, j, k));
            const double vze = dz * (PQ(3, j, k + 1) - PQ(3, j, k));
            const double wze = dz * (PQ(4, j, k + 1) - PQ(4, j, k));
            const double b1 = ETA(1, ip, j, k + 1) + ETA(1, ip, j, k);
            const double b2 = ETA(2, ip, j, k + 1) + ETA(2, ip, j, k);
            const double b3 = ETA(3, ip, j, k + 1) + ETA(3, ip, j, k);
            const double c1 = ZET(1, ip, j, k + 1) + ZET(1, ip, j, k);
            const double c2 = ZET(2, ip, j, k + 1) + ZET(2, ip, j, k);
            const double c3 = ZET(3, ip, j, k + 1) + ZET(3, ip, j, k);
            const double uu1 = b1 * uet + c1 * uze, vv1 = b1 * vet + c1 * vze, ww1 = b1 * wet + c1 * wze;
            const double uu2 = b2 * uet + c2 * uze, vv2 = b2 * vet + c2 * vze, ww2 = b2 * wet + c2 * wze;
            const double uu3 = b3 * uet + c3 * uze, vv3 = b3 * vet + c3 * vze, ww3 = b3 * wet + c3 * wze;
            const double c32 = c1 * b1 + c2 * b2 + c3 * b3;
            const double c33 = c1 * c1 + c2 * c2 + c3 * c3;
            const double vtj = one_fourth * (ret[P2(j, k + 1)] + ret[P2(j, k)]) / (AJ(ip, j, k + 1) + AJ(ip, j, k));
            fv[3 * P2(j, k) + 0] = vtj * (c1 * uu1 + c2 * vv1 + c3 * ww1 + c32 * uet + c33 * uze);
            fv[3 * P2(j, k) + 1] = vtj * (c1 * uu2 + c2 * vv2 + c3 * ww2 + c32 * vet + c33 * vze);
            fv[3 * P2(j, k) + 2] = vtj * (c1 * uu3 + c2 * vv3 + c3 * ww3 + c32 * wet + c33 * wze);
        }
    for (k = ks; k <= ke; k++)
        for (j = js; j <= je; j++)
            for (c = 0; c < 3; c++)
                visc[3 * P2(j, k) + c] = visc[3 * P2(j, k) + c] + dz * (fv[3 * P2(j, k) + c] - fv[3 * P2(j, k - 1) + c]);
    /* brhs_diss_p */
    const double desq = de * de, dzsq = dz * dz;
    for (k = ks; k <= ke; k++)
        for (j = js; j <= je; j++) {
            const double g2 = ETA(1, ip, j, k) * ETA(1, ip, j, k) + ETA(2, ip, j, k) * ETA(2, ip, j, k) +
                              ETA(3, ip, j, k) * ETA(3, ip, j, k);
            const double tmp1 = g2 * DTAU(ip, j, k) / AJ(ip, j, k);
            const double tmp2 = desq * (PQ(1, j + 1, k) - 2.0 * PQ(1, j, k) + PQ(1, j - 1, k));
            aux[2 * P2(j, k) + 0] = tmp1 * tmp2;
        }
    for (k = ks; k <= ke; k++)
        for (j = js; j <= je; j++) {
            const double g3 = ZET(1, ip, j, k) * ZET(1, ip, j, k) + ZET(2, ip, j, k) * ZET(2, ip, j, k) +
                              ZET(3, ip, j, k) * ZET(3, ip, j, k);
            const double tmp1 = g3 * DTAU(ip, j, k) / AJ(ip, j, k);
            const double tmp2 = dzsq * (PQ(1, j, k + 1) - 2.0 * PQ(1, j, k) + PQ(1, j, k - 1));
            aux[2 * P2(j, k) + 1] = tmp1 * tmp2;
        }
    for (k = 1; k <= nk; k++) {
        aux[2 * P2(js - 1, k)] = aux[2 * P2(js, k)];
        aux[2 * P2(je + 1, k)] = aux[2 * P2(je, k)];
    }
    for (j = 1; j <= nj; j++) {
        aux[2 * P2(j, ks - 1) + 1] = aux[2 * P2(j, ks) + 1];
        aux[2 * P2(j, ke + 1) + 1] = aux[2 * P2(j, ke) + 1];
    }
    for (k = ks; k <= ke; k++)
        for (j = js; j <= je; j++)
            diss[4 * P2(j, k)] = o->p.pdiss_coef * ((aux[2 * P2(j + 1, k)] - 2.0 * aux[2 * P2(j, k)] + aux[2 * P2(j - 1, k)]) +
                                                    (aux[2 * P2(j, k + 1) + 1] - 2.0 * aux[2 * P2(j, k) + 1] + aux[2 * P2(j, k - 1) + 1]));
    /* brhs_flux_sans_convec */
    const double de2 = pt5 * de, dz2 = pt5 * dz;
    for (k = 1; k <= nk; k++)
        for (j = 1; j <= nj; j++) pp[P2(j, k)] = PQ(1, j, k) / AJ(ip, j, k);
    for (k = ks; k <= ke; k++)
        for (j = js; j <= je; j++)
            for (c = 1; c <= 3; c++)
                RH(c + 1, ip, j, k) = de2 * (pp[P2(j + 1, k)] * ETA(c, ip, j + 1, k) - pp[P2(j - 1, k)] * ETA(c, ip, j - 1, k));
    for (k = ks; k <= ke; k++)
        for (j = js; j <= je; j++)
            for (c = 1; c <= 3; c++)
                RH(c + 1, ip, j, k) = RH(c + 1, ip, j, k) +
                                      dz2 * (pp[P2(j, k + 1)] * ZET(c, ip, j, k + 1) - pp[P2(j, k - 1)] * ZET(c, ip, j, k - 1));
    for (k = ks; k <= ke; k++)
        for (j = js; j <= je; j++) {
            const double dudx = 0.0;
            const double dvdy = de2 * (UCN(2, ip, j + 1, k) - UCN(2, ip, j - 1, k));
            const double dwdz = dz2 * (UCN(3, ip, j, k + 1) - UCN(3, ip, j, k - 1));
            RH(1, ip, j, k) = dudx + dvdy + dwdz;
        }
    /* brhs_convec_quick(2) */
    const double coef = 0.125;
    for (int dir = 2; dir <= 3; dir++) {
        const double h = dir == 2 ? de : dz;
        const int dj = dir == 2, dk = dir == 3;
        for (k = 1; k <= nk; k++)
            for (j = 1; j <= nj; j++) {
                const double ucon = one_half * UCN(dir, ip, j, k);
                up[P2(j, k)] = ucon + fabs(ucon);
                um[P2(j, k)] = ucon - fabs(ucon);
            }
        const int sta = dir == 2 ? js + 1 : ks + 1, end = dir == 2 ? je - 1 : ke - 1;
        for (k = ks; k <= ke; k++)
            for (j = js; j <= je; j++) {
                const int p = dir == 2 ? j : k;
                double* r = rd + 3 * P2(j, k);
                const double upv = up[P2(j, k)], umv = um[P2(j, k)];
                for (c = 0; c < 3; c++) {
                    const int m = c + 2;
                    if (p >= sta && p <= end) {
                        r[c] = coef * h *
                               (upv * (3.0 * PQ(m, j + dj, k + dk) + 3.0 * PQ(m, j, k) - 7.0 * PQ(m, j - dj, k - dk) +
                                       PQ(m, j - 2 * dj, k - 2 * dk)) +
                                umv * (-3.0 * PQ(m, j - dj, k - dk) - 3.0 * PQ(m, j, k) + 7.0 * PQ(m, j + dj, k + dk) -
                                       PQ(m, j + 2 * dj, k + 2 * dk)));
                    } else {
                        r[c] = h * (upv * (PQ(m, j, k) - PQ(m, j - dj, k - dk)) + umv * (PQ(m, j + dj, k + dk) - PQ(m, j, k)));
                    }
                }
            }
        for (k = ks; k <= ke; k++)
            for (j = js; j <= je; j++)
                for (c = 0; c < 3; c++) RH(c + 2, ip, j, k) = RH(c + 2, ip, j, k) + rd[3 * P2(j, k) + c];
    }
    /* brhs_unst_visc_diss (diss(2:4) = 0, set by brhs_convec_quick) */
    for (k = ks; k <= ke; k++)
        for (j = js; j <= je; j++) {
            const double dtaj = 2.0 * o->p.delti * AJ(ip, j, k);
            RH(1, ip, j, k) = RH(1, ip, j, k) + diss[4 * P2(j, k)];
            for (c = 1; c <= 3; c++) {
                const int m = c + 1;
                const double dqdt = o->e_source * (3.0 * PQ(m, j, k) - 4.0 * QN(m, ip, j, k) + QNM1(m, ip, j, k)) / dtaj;
                RH(m, ip, j, k) = RH(m, ip, j, k) + 0.0 + dqdt - visc[3 * P2(j, k) + c - 1];
            }
        }
#undef PQ
#undef P2
    free(visc); free(diss); free(fv); free(ret); free(aux); free(pp); free(up); free(um); free(rd);
}

static void nonreflect_plane(daf_t* d, int ip, int inout) {
    DLV;
    const int nb = inout == 1 ? ip + 1 : ip - 1;  /* the interior neighbour plane */
    for (int k = 2; k <= nk - 1; k++)
        for (int j = 2; j <= nj - 1; j++) {
            const double u = Q(2, ip, j, k), v = Q(3, ip, j, k), w = Q(4, ip, j, k);
            const double ucon = UCN(1, ip, j, k) * AJ(ip, j, k);
            const double c1 = CSI(1, ip, j, k), c2 = CSI(2, ip, j, k), c3 = CSI(3, ip, j, k);
            const double gjj = c1 * c1 + c2 * c2 + c3 * c3;
            const double sg = sqrt(gjj);
            const double cj = sqrt(ucon * ucon + gjj);
            double dq[4];
            for (int m = 1; m <= 4; m++)
                dq[m - 1] = inout == 1 ? (Q(m, nb, j, k) - Q(m, ip, j, k)) : (Q(m, ip, j, k) - Q(m, nb, j, k));
            /* calc_L_general */
            const double lam1 = ucon - cj, lam2 = ucon + cj, lam3 = ucon, lam4 = ucon;
            const double kj = c1 + c2 + c3;
            const double d1 = (c2 - c1 + lam1 * (v - u)) / kj;
            const double d2 = (c2 - c1 + lam2 * (v - u)) / kj;
            const double d3 = (c3 - c1 + lam1 * (w - u)) / kj;
            const double d4 = (c3 - c1 + lam2 * (w - u)) / kj;
            const double d5 = (c2 - c1 + lam3 * (v - u)) / kj * 2.0;
            const double d6 = (c3 - c1 + lam4 * (w - u)) / kj * 2.0;
            const double t2 = 2.0 * cj * cj / kj;
            double si[4][4] = {{-lam2, c1, c2, c3},
                               {-lam1, c1, c2, c3},
                               {lam2 * d1 + lam1 * d2, -t2 - c1 * d5, t2 - c2 * d5, -c3 * d5},
                               {lam2 * d3 + lam1 * d4, -t2 - c1 * d6, -c2 * d6, t2 - c3 * d6}};
            const double den = 2.0 * cj * cj;
            double lg[4];
            for (int r = 0; r < 4; r++) {
                double acc = 0.0;
                for (int cc = 0; cc < 4; cc++) acc = acc + si[r][cc] * sg / den * dq[cc];
                lg[r] = acc;
            }
            lg[0] = lam1 * lg[0];
            lg[1] = lam2 * lg[1];
            lg[2] = lam3 * lg[2];
            lg[3] = lam4 * lg[3];
            double L1, L2, L3, L4;
            if (inout == 1) { /* calc_L_min */
                L1 = lg[0]; L2 = 0.0;
                if (ucon < 0.0) { L3 = lg[2]; L4 = lg[3]; } else { L3 = 0.0; L4 = 0.0; }
            } else {          /* calc_L_max */
                L1 = 0.0; L2 = lg[1];
                if (ucon < 0.0) { L3 = 0.0; L4 = 0.0; } else { L3 = lg[2]; L4 = lg[3]; }
            }
            const double rh1 = (-cj * L1 + cj * L2) / sg;
            const double rh2 = ((c1 + u * lam1) * L1 + (c1 + u * lam2) * L2 - c2 * L3 - c3 * L4) / sg;
            const double rh3 = ((c2 + v * lam1) * L1 + (c2 + v * lam2) * L2 + (c1 + c3) * L3 - c3 * L4) / sg;
            const double rh4 = ((c3 + w * lam1) * L1 + (c3 + w * lam2) * L2 - c2 * L3 + (c1 + c2) * L4) / sg;
            RH(1, ip, j, k) = RH(1, ip, j, k) + rh1 / AJ(ip, j, k);
            RH(2, ip, j, k) = RH(2, ip, j, k) + rh2 / AJ(ip, j, k);
            RH(3, ip, j, k) = RH(3, ip, j, k) + rh3 / AJ(ip, j, k);
            RH(4, ip, j, k) = RH(4, ip, j, k) + rh4 / AJ(ip, j, k);
        }
}

/* rhs_modal_matrices (src/rhs_modal_matrices.F90:1-270) */
static void matmul4(const double a[4][4], const double b[4][4], double c[4][4]) {
    int r, s, t;
    for (r = 0; r < 4; r++)
        for (s = 0; s < 4; s++) {
            double acc = 0.0;
            for (t = 0; t < 4; t++) acc = acc + a[r][t] * b[t][s];
            c[r][s] = acc;
        }
}

static void rhs_modal_matrices(daf_t* d) {
    DLV;
    int i, j, k, r, c;
    const double es = d->o->e_source, delti = d->o->p.delti, beta = d->o->p.beta;
    for (k = 1; k <= nk; k++)
        for (j = 1; j <= nj; j++)
            for (i = 1; i <= ni; i++) {
                double csi1 = CSI(1, i, j, k), csi2 = CSI(2, i, j, k), csi3 = CSI(3, i, j, k);
                double eta1 = ETA(1, i, j, k), eta2 = ETA(2, i, j, k), eta3 = ETA(3, i, j, k);
                double zet1 = ZET(1, i, j, k), zet2 = ZET(2, i, j, k), zet3 = ZET(3, i, j, k);
                double csi1s = csi1 * csi1, csi2s = csi2 * csi2, csi3s = csi3 * csi3;
                double eta1s = eta1 * eta1, eta2s = eta2 * eta2, eta3s = eta3 * eta3;
                double zet1s = zet1 * zet1, zet2s = zet2 * zet2, zet3s = zet3 * zet3;
                double g1 = csi1s + csi2s + csi3s, g2 = eta1s + eta2s + eta3s, g3 = zet1s + zet2s + zet3s;
                double sg1 = sqrt(g1), sg2 = sqrt(g2), sg3 = sqrt(g3);
                double alpha = 1.0 + es * onept5 * DTAU(i, j, k) / delti;
                double sigma = sqrt(beta * alpha), theta = sqrt(beta / alpha);
                double scsi, seta, szet, dcsi, deta, dzet, rc1, rc2, rc3, re1, re2, re3, rz1, rz2, rz3, ssgd, sgd2;
                double ma[4][4], mb[4][4], mbi[4][4], mci[4][4], mai[4][4], mc[4][4], rp[4][4];
                SPR(1, i, j, k) = theta * sg1;
                SPR(2, i, j, k) = theta * sg2;
                SPR(3, i, j, k) = theta * sg3;
                scsi = sqrt(2.0 * g1 + (csi2 - csi3) * (csi2 - csi3));
                seta = sqrt(2.0 * g2 + (eta2 - eta3) * (eta2 - eta3));
                szet = sqrt(2.0 * g3 + (zet2 - zet3) * (zet2 - zet3));
                dcsi = g1 + 2.0 * csi2 * csi3 + csi1 * csi1;
                deta = g2 + 2.0 * eta2 * eta3 + eta1 * eta1;
                dzet = g3 + 2.0 * zet2 * zet3 + zet1 * zet1;
                rc1 = csi2 + csi3; rc2 = csi2 - csi3; rc3 = csi2s - csi3s;
                re1 = eta2 + eta3; re2 = eta2 - eta3; re3 = eta2s - eta3s;
                rz1 = zet2 + zet3; rz2 = zet2 - zet3; rz3 = zet2s - zet3s;
                /* Ma */
                ma[0][0] = 0.0; ma[0][1] = 0.0; ma[0][2] = sigma; ma[0][3] = sigma;
                ma[1][0] = -(csi2 + csi3) / scsi; ma[1][1] = csi1 * (csi3 - csi2) / (scsi * sg1); ma[1][2] = csi1 / sg1; ma[1][3] = -csi1 / sg1;
                ma[2][0] = csi1 / scsi; ma[2][1] = (g1 + csi2 * (csi3 - csi2)) / (scsi * sg1); ma[2][2] = csi2 / sg1; ma[2][3] = -csi2 / sg1;
                ma[3][0] = csi1 / scsi; ma[3][1] = -(g1 - csi3 * (csi3 - csi2)) / (scsi * sg1); ma[3][2] = csi3 / sg1; ma[3][3] = -csi3 / sg1;
                /* Mai */
                mai[0][0] = 0.0; mai[0][1] = -scsi * (csi2 + csi3) / dcsi; mai[0][2] = scsi * csi1 / dcsi; mai[0][3] = mai[0][2];
                ssgd = scsi / sg1 / dcsi;
                sgd2 = 1.0 / sg1 / dcsi / 2.0;
                mai[1][0] = 0.0; mai[1][1] = csi1 * (csi3 - csi2) * ssgd; mai[1][2] = (csi2 * csi3 + csi1s + csi3s) * ssgd; mai[1][3] = -(csi2 * csi3 + csi1s + csi2s) * ssgd;
                mai[2][0] = pt5 / sigma; mai[2][1] = csi1 * (2.0 * g1 + 2.0 * csi2 * csi3 - csi2s - csi3s) * sgd2;
                mai[2][2] = (g1 * rc1 + csi1s * rc2 + csi3 * rc3) * sgd2; mai[2][3] = (g1 * rc1 - csi1s * rc2 - csi2 * rc3) * sgd2;
                mai[3][0] = pt5 / sigma; mai[3][1] = -mai[2][1]; mai[3][2] = -mai[2][2]; mai[3][3] = -mai[2][3];
                /* Mb */
                mb[0][0] = 0.0; mb[0][1] = 0.0; mb[0][2] = sigma; mb[0][3] = sigma;
                mb[1][0] = -(eta2 + eta3) / seta; mb[1][1] = eta1 * (eta3 - eta2) / (seta * sg2); mb[1][2] = eta1 / sg2; mb[1][3] = -eta1 / sg2;
                mb[2][0] = eta1 / seta; mb[2][1] = (g2 + eta2 * (eta3 - eta2)) / (seta * sg2); mb[2][2] = eta2 / sg2; mb[2][3] = -eta2 / sg2;
                mb[3][0] = eta1 / seta; mb[3][1] = -(g2 - eta3 * (eta3 - eta2)) / (seta * sg2); mb[3][2] = eta3 / sg2; mb[3][3] = -eta3 / sg2;
                /* Mbi */
                mbi[0][0] = 0.0; mbi[0][1] = -seta * (eta2 + eta3) / deta; mbi[0][2] = seta * eta1 / deta; mbi[0][3] = mbi[0][2];
                ssgd = seta / sg2 / deta;
                sgd2 = 1.0 / sg2 / deta / 2.0;
                mbi[1][0] = 0.0; mbi[1][1] = eta1 * (eta3 - eta2) * ssgd; mbi[1][2] = (eta2 * eta3 + eta1s + eta3s) * ssgd; mbi[1][3] = -(eta2 * eta3 + eta1s + eta2s) * ssgd;
                mbi[2][0] = pt5 / sigma; mbi[2][1] = eta1 * (2.0 * g2 + 2.0 * eta2 * eta3 - eta2s - eta3s) * sgd2;
                mbi[2][2] = (g2 * re1 + eta1s * re2 + eta3 * re3) * sgd2; mbi[2][3] = (g2 * re1 - eta1s * re2 - eta2 * re3) * sgd2;
                mbi[3][0] = pt5 / sigma; mbi[3][1] = -mbi[2][1]; mbi[3][2] = -mbi[2][2]; mbi[3][3] = -mbi[2][3];
                /* Mc */
                mc[0][0] = 0.0; mc[0][1] = 0.0; mc[0][2] = sigma; mc[0][3] = sigma;
                mc[1][0] = -(zet2 + zet3) / szet; mc[1][1] = zet1 * (zet3 - zet2) / (szet * sg3); mc[1][2] = zet1 / sg3; mc[1][3] = -zet1 / sg3;
                mc[2][0] = zet1 / szet; mc[2][1] = (g3 + zet2 * (zet3 - zet2)) / (szet * sg3); mc[2][2] = zet2 / sg3; mc[2][3] = -zet2 / sg3;
                mc[3][0] = zet1 / szet; mc[3][1] = -(g3 - zet3 * (zet3 - zet2)) / (szet * sg3); mc[3][2] = zet3 / sg3; mc[3][3] = -zet3 / sg3;
                /* Mci */
                mci[0][0] = 0.0; mci[0][1] = -szet * (zet2 + zet3) / dzet; mci[0][2] = szet * zet1 / dzet; mci[0][3] = mci[0][2];
                ssgd = szet / sg3 / dzet;
                sgd2 = 1.0 / sg3 / dzet / 2.0;
                mci[1][0] = 0.0; mci[1][1] = zet1 * (zet3 - zet2) * ssgd; mci[1][2] = (zet2 * zet3 + zet1s + zet3s) * ssgd; mci[1][3] = -(zet2 * zet3 + zet1s + zet2s) * ssgd;
                mci[2][0] = pt5 / sigma; mci[2][1] = zet1 * (2.0 * g3 + 2.0 * zet2 * zet3 - zet2s - zet3s) * sgd2;
                mci[2][2] = (g3 * rz1 + zet1s * rz2 + zet3 * rz3) * sgd2; mci[2][3] = (g3 * rz1 - zet1s * rz2 - zet2 * rz3) * sgd2;
                mci[3][0] = pt5 / sigma; mci[3][1] = -mci[2][1]; mci[3][2] = -mci[2][2]; mci[3][3] = -mci[2][3];
                for (r = 0; r < 4; r++)
                    for (c = 0; c < 4; c++) {
                        MAT(mai, r + 1, c + 1, i, j, k) = mai[r][c];
                        MAT(mc, r + 1, c + 1, i, j, k) = mc[r][c];
                    }
                matmul4(mbi, ma, rp);
                for (r = 0; r < 4; r++)
                    for (c = 0; c < 4; c++) MAT(n1i, r + 1, c + 1, i, j, k) = rp[r][c];
                matmul4(mci, mb, rp);
                for (r = 0; r < 4; r++)
                    for (c = 0; c < 4; c++) MAT(n2i, r + 1, c + 1, i, j, k) = rp[r][c];
            }
}

static void matvec_node(const double* a /* 16, row-major */, double* v /* 4 */) {
    double rp[4];
    int r, c;
    for (r = 0; r < 4; r++) {
        double acc = 0.0;
        for (c = 0; c < 4; c++) acc = acc + a[4 * r + c] * v[c];
        rp[r] = acc;
    }
    for (r = 0; r < 4; r++) v[r] = rp[r];
}

/* one tridiagonal line solve of the DAF: builds the 4 scalar systems of
 * rhs_diag_solver for a line in direction dir and solves them with SGTSV */
static void daf_line(daf_t* d, int dir, int i0, int j0, int k0, int len, const double* brh1, const double* brh2,
                     int use_b1, int use_b2, double* aw, double* ap, double* ae, double* aq, int ext) {
    DLV;
    int s, m;
    const double h = dir == 1 ? d->dc : dir == 2 ? d->de : d->dz;
    const double h2 = pt5 * h, hsq = h * h;
    const double epd = d->o->p.ep[0][dir - 1];
    const double lambda[4] = {0.0, 0.0, 1.0, -1.0};
    const int di = dir == 1, dj = dir == 2, dk = dir == 3;
    (void)nk;
    for (s = 0; s < len - ext; s++) {
        int i = i0 + s * di, j = j0 + s * dj, k = k0 + s * dk;
        double dtc = 1.0, tmp, dds;
        if (DTAU(i, j, k) == 0.0) dtc = 0.0;
        tmp = DTAU(i, j, k) * epd * SPR(dir, i, j, k) * hsq;
        for (m = 0; m < 4; m++) ap[m * len + s] = 1.0 + 2.0 * tmp;
        dds = dtc * DTAU(i - di, j - dj, k - dk) * h2 * SPR(dir, i - di, j - dj, k - dk);
        aw[0 * len + s] = -tmp;
        aw[1 * len + s] = -tmp;
        aw[2 * len + s] = -dds * lambda[2] - tmp;
        aw[3 * len + s] = -dds * lambda[3] - tmp;
        dds = dtc * DTAU(i + di, j + dj, k + dk) * h2 * SPR(dir, i + di, j + dj, k + dk);
        ae[0 * len + s] = -tmp;
        ae[1 * len + s] = -tmp;
        ae[2 * len + s] = dds * lambda[2] - tmp;
        ae[3 * len + s] = dds * lambda[3] - tmp;
        for (m = 0; m < 4; m++) aq[m * len + s] = RH(m + 1, i, j, k);
    }
    if (ext) {
        /* the characteristic-outlet plane i = idend closes the i line
         * (rhs_diag_solver.F90:207-221): upwind coupling of the +1
         * eigenvalue only, no smoothing */
        const int i = i0 + (len - 1);
        double dds = DTAU(i, j0, k0) * h * SPR(1, i, j0, k0);
        s = len - 1;
        ap[0 * len + s] = 1.0;
        ap[1 * len + s] = 1.0;
        ap[2 * len + s] = 1.0 + dds * lambda[2];
        ap[3 * len + s] = 1.0;
        dds = DTAU(i - 1, j0, k0) * h * SPR(1, i - 1, j0, k0);
        aw[0 * len + s] = 0.0;
        aw[1 * len + s] = 0.0;
        aw[2 * len + s] = -dds * lambda[2];
        aw[3 * len + s] = 0.0;
        for (m = 0; m < 4; m++) ae[m * len + s] = 0.0;
        for (m = 0; m < 4; m++) aq[m * len + s] = RH(m + 1, i, j0, k0);
    }
    if (use_b1)
        for (m = 0; m < 4; m++) aq[m * len + 0] = aq[m * len + 0] - aw[m * len + 0] * brh1[m];
    if (use_b2)
        for (m = 0; m < 4; m++) aq[m * len + len - 1] = aq[m * len + len - 1] - ae[m * len + len - 1] * brh2[m];
    for (m = 0; m < 4; m++)
        or_gtsv(len, 1, aw + m * len + 1, ap + m * len, ae + m * len, aq + m * len, len);
    for (s = 0; s < len; s++) {
        int i = i0 + s * di, j = j0 + s * dj, k = k0 + s * dk;
        for (m = 0; m < 4; m++) RH(m + 1, i, j, k) = aq[m * len + s];
    }
}

/* rhs_diag_solver (src/rhs_diag_solver.F90:2-636) */
static void rhs_diag_solver(daf_t* d) {
    DLV;
    const or_state* o = d->o;
    int i, j, k, m, c;
    const int n = d->n;
    const int is = d->is, ie = d->ie, js = d->js, je = d->je, ks = d->ks, ke = d->ke, idend = d->idend;
    const double alpha_mult = o->e_source * one_pt_five / o->p.delti;
    int maxlen = ni > nj ? ni : nj;
    double *aw, *ap, *ae, *aq, *brh;
    const int *bt = o->p.btype;
    if (nk > maxlen) maxlen = nk;
    aw = zalloc(4 * maxlen); ap = zalloc(4 * maxlen); ae = zalloc(4 * maxlen); aq = zalloc(4 * maxlen);
    /* brh(1:4, 1:2, line); the j and k sweeps index lines by (i, k) and
     * (i, j) with i up to idend = ni on a characteristic outlet */
    brh = zalloc(8L * (long)(maxlen + 1) * (maxlen + 1));
#define BRH(m, side, a, b, na) brh[((long)((b)) * (na) + (a)) * 8 + ((side)-1) * 4 + (m)-1]
    /* rh = S (-dtau J rh) (:56-72) */
    for (k = ks; k <= ke; k++)
        for (j = js; j <= je; j++)
            for (i = is; i <= idend; i++) {
                double alpha = 1.0 + alpha_mult * DTAU(i, j, k);
                double r[4];
                for (m = 1; m <= 4; m++) r[m - 1] = -DTAU(i, j, k) * RH(m, i, j, k) * AJ(i, j, k);
                RH(1, i, j, k) = o->p.beta * r[0];
                for (m = 2; m <= 4; m++) RH(m, i, j, k) = 1.0 / alpha * r[m - 1];
            }
    /* ---- csi sweep ---- */
    memset(brh, 0, 8L * (maxlen + 1) * (maxlen + 1) * sizeof(double));
    if (n == 1 && bt[0] != 5)
        for (k = ks; k <= ke; k++)
            for (j = js; j <= je; j++) {
                for (m = 1; m <= 4; m++)
                    BRH(m, 1, j, k, nj) = o->sa[0][m - 1] * RH(m, is, j, k) + o->sb[0][m - 1] * RH(m, is + 1, j, k);
                /* sa(:,1) has 5 entries, rh(:,i,j,k) 4 (:97): the fifth
                 * product reads rh(1) of the next i node and is stored in
                 * brh(1,2,j,k) */
                BRH(1, 2, j, k, nj) = o->sa[0][4] * RH(1, is + 1, j, k) + o->sb[0][4] * RH(1, is + 2, j, k);
            }
    for (k = ks; k <= ke; k++)
        for (j = js; j <= je; j++)
            for (i = is; i <= idend; i++) matvec_node(&MAT(mai, 1, 1, i, j, k), &RH(1, i, j, k));
    if (n == 1 && bt[0] != 5)
        for (k = ks; k <= ke; k++)
            for (j = js; j <= je; j++) matvec_node(&MAT(mai, 1, 1, is - 1, j, k), &BRH(1, 1, j, k, nj));
    for (k = ks; k <= ke; k++)
        for (j = js; j <= je; j++) {
            /* front-side brh is built only under myfront /= MPI_PROC_NULL (:102),
             * so on the domain boundary it stays zero when consumed (:202-205) */
            daf_line(d, 1, is, j, k, idend - is + 1, &BRH(1, 1, j, k, nj), &BRH(1, 2, j, k, nj),
                     n == 1 && bt[0] != 5, n == 1 && bt[1] != 5, aw, ap, ae, aq, idend > ie);
        }
    /* ---- eta sweep ---- */
    memset(brh, 0, 8L * (maxlen + 1) * (maxlen + 1) * sizeof(double));
    if (n == 1) {
        if (bt[2] != 5)
            for (k = ks; k <= ke; k++)
                for (i = is; i <= idend; i++) {
                    for (m = 1; m <= 4; m++)
                        BRH(m, 1, i, k, ni) = o->sa[2][m - 1] * RH(m, i, js, k) + o->sb[2][m - 1] * RH(m, i, js + 1, k);
                    /* fifth product of the 5 x 4 non-conforming expression (:286) */
                    BRH(1, 2, i, k, ni) = o->sa[2][4] * RH(1, i + 1, js, k) + o->sb[2][4] * RH(1, i + 1, js + 1, k);
                }
        if (bt[3] != 5)
            for (k = ks; k <= ke; k++)
                for (i = is; i <= idend; i++)
                    for (m = 1; m <= 4; m++)
                        BRH(m, 2, i, k, ni) = o->sa[3][m - 1] * RH(m, i, je, k) + o->sb[3][m - 1] * RH(m, i, je - 1, k);
    }
    for (k = ks; k <= ke; k++)
        for (j = js; j <= je; j++)
            for (i = is; i <= idend; i++) matvec_node(&MAT(n1i, 1, 1, i, j, k), &RH(1, i, j, k));
    if (n == 1) {
        if (bt[2] != 5)
            for (k = ks; k <= ke; k++)
                for (i = is; i <= idend; i++) matvec_node(&MAT(n1i, 1, 1, i, 1, k), &BRH(1, 1, i, k, ni));
        if (bt[3] != 5)
            for (k = ks; k <= ke; k++)
                for (i = is; i <= idend; i++) matvec_node(&MAT(n1i, 1, 1, i, je + 1, k), &BRH(1, 2, i, k, ni));
    }
    for (k = ks; k <= ke; k++)
        for (i = is; i <= idend; i++)
            daf_line(d, 2, i, js, k, je - js + 1, &BRH(1, 1, i, k, ni), &BRH(1, 2, i, k, ni),
                     n == 1 && bt[2] != 5, n == 1 && bt[3] != 5, aw, ap, ae, aq, 0);
    /* ---- zet sweep ---- */
    memset(brh, 0, 8L * (maxlen + 1) * (maxlen + 1) * sizeof(double));
    if (n == 1) {
        if (bt[4] != 5)
            for (j = js; j <= je; j++)
                for (i = is; i <= idend; i++) {
                    for (m = 1; m <= 4; m++)
                        BRH(m, 1, i, j, ni) = o->sa[4][m - 1] * RH(m, i, j, ks) + o->sb[4][m - 1] * RH(m, i, j, ks + 1);
                    /* fifth product of the 5 x 4 non-conforming expression (:462) */
                    BRH(1, 2, i, j, ni) = o->sa[4][4] * RH(1, i + 1, j, ks) + o->sb[4][4] * RH(1, i + 1, j, ks + 1);
                }
        /* built under btype(6) /= 6 (:467) but consumed under /= 5 (:504,571) */
        if (bt[5] != 6)
            for (j = js; j <= je; j++)
                for (i = is; i <= idend; i++)
                    for (m = 1; m <= 4; m++)
                        BRH(m, 2, i, j, ni) = o->sa[5][m - 1] * RH(m, i, j, ke) + o->sb[5][m - 1] * RH(m, i, j, ke - 1);
    }
    for (k = ks; k <= ke; k++)
        for (j = js; j <= je; j++)
            for (i = is; i <= idend; i++) matvec_node(&MAT(n2i, 1, 1, i, j, k), &RH(1, i, j, k));
    if (n == 1) {
        if (bt[4] != 5)
            for (j = js; j <= je; j++)
                for (i = is; i <= idend; i++) matvec_node(&MAT(n2i, 1, 1, i, j, 1), &BRH(1, 1, i, j, ni));
        if (bt[5] != 5)
            for (j = js; j <= je; j++)
                for (i = is; i <= idend; i++) matvec_node(&MAT(n2i, 1, 1, i, j, ke + 1), &BRH(1, 2, i, j, ni));
    }
    for (j = js; j <= je; j++)
        for (i = is; i <= idend; i++)
            daf_line(d, 3, i, j, ks, ke - ks + 1, &BRH(1, 1, i, j, ni), &BRH(1, 2, i, j, ni),
                     n == 1 && bt[4] != 5, n == 1 && bt[5] != 5, aw, ap, ae, aq, 0);
    /* dQ = Mc dQ*** (:624-633) */
    for (k = ks; k <= ke; k++)
        for (j = js; j <= je; j++)
            for (i = is; i <= idend; i++) matvec_node(&MAT(mc, 1, 1, i, j, k), &RH(1, i, j, k));
    (void)c;
#undef BRH
    free(aw); free(ap); free(ae); free(aq); free(brh);
}

/* ------------------------------------------------------------------ */
/* Actuator-line model: rhs_lines (src/rhs_lines.F90:47-483), an        */
/* internal procedure of solver_daf run on every level after the rh     */
/* boundary zeroing (solver_daf.F90:378-381).  Node indices of the      */
/* tables are fine-level indices; on level n they are used as they are, */
/* against level n's interior range and arrays, as the reference does.  */
/* ------------------------------------------------------------------ */
/* Cl, Cd of polar P at angle of attack aoa: the first pair of consecutive
 * points with angle1 < aoa <= angle2, linear in between
 * (rhs_lines.F90:232-265).  Outside the polar the reference keeps the
 * previous element's values; here the end point's values are taken. */
static void polar_lookup(const desalm_polar* P, double aoa, double* cl, double* cd) {
    int idx;
    if (aoa <= P->angle[0]) {
        *cl = P->cl[0];
        *cd = P->cd[0];
        return;
    }
    for (idx = 0; idx + 1 < P->n; idx++) {
        const double a1 = P->angle[idx], a2 = P->angle[idx + 1];
        if (aoa <= a1) break;
        if (aoa <= a2 && aoa >= a1) {
            double m = (P->cl[idx + 1] - P->cl[idx]) / (a2 - a1);
            *cl = m * (aoa - a1) + P->cl[idx];
            m = (P->cd[idx + 1] - P->cd[idx]) / (a2 - a1);
            *cd = m * (aoa - a1) + P->cd[idx];
            return;
        }
    }
    *cl = P->cl[P->n - 1];
    *cd = P->cd[P->n - 1];
}

static void rhs_lines(daf_t* d) {
    const or_state* o = d->o;
    if (o->n_tur == 0) return;
    const int ni = d->ni, nj = d->nj;
    const int jmg = o->jmg[1], kmg = o->kmg[1];
    const double pi = 3.141592653589793;
    const double rho = 1178.62;
    const double L = o->length_scale;
    const double vel_scale = o->p.ren * 10E-7 / L;
#define QL(m, i, j, k) d->q[5 * IX(i, j, k) + (m)-1]
    for (int t = 0; t < o->n_tur; t++) {
        const desalm_turbine* T = &o->tur[t];
        const int ne = T->nelems, na = T->nangles;
        double* fl = (double*)calloc((size_t)T->nblades * ne * 6, sizeof(double)); /* forces_lines = 0 */
        const double omega = (o->vel_infty * T->tsr) / T->tur_rad;
        for (int bld = 1; bld <= T->nblades; bld++) {
            double angle_aux = T->angle + (bld - 1) * 2 * pi / T->nblades;
            if (angle_aux > 2 * pi) angle_aux = angle_aux - 2 * pi;
            const int nang = (int)lround(angle_aux / T->dtheta_interp) + 1;
            for (int elem = 1; elem <= ne; elem++) {
                const long e = (elem - 1) + (long)ne * (nang - 1);
                const int i = T->node_indexes[e], j = T->node_indexes[e + (long)ne * na],
                          k = T->node_indexes[e + 2L * ne * na];
                const double* ps = T->node_indexes_pos;
#define NP(c) ps[e + (long)ne * na * ((c)-1)]
                const double xf = NP(1), yf = NP(2), zf = NP(3);
                const double x000 = NP(4), y000 = NP(5), z000 = NP(6), x100 = NP(7), y100 = NP(8);
                const double x110 = NP(10), y110 = NP(11), z001 = NP(18), x001 = NP(16);
                const double x101 = NP(19), y101 = NP(20), y111 = NP(23);
#undef NP
                (void)z000;
                if (!(i >= d->is && i <= d->ie && j >= d->js && j <= d->je && k >= d->ks && k <= d->ke)) continue;
                const double* dd = T->disk_dat;
                const long jk = (j - 1) + (long)jmg * (k - 1), ck = (long)jmg * kmg;
                const double radius = dd[jk] * L;
                const double chord = dd[jk + 4 * ck], twist = dd[jk + 5 * ck], inter = dd[jk + 6 * ck];
                double uvw[3];
                for (int m = 2; m <= 4; m++) {
                    const double down = 1 / ((x100 - x000) * (y110 - y100)) *
                                        ((x100 - xf) * ((y110 - yf) * (QL(m, i, j, k) + QL(m, i, j + 1, k))) +
                                         (xf - x000) * ((yf - y100) * (QL(m, i + 1, j, k) + QL(m, i + 1, j + 1, k))));
                    const double up = 1 / ((x101 - x001) * (y111 - y101)) *
                                      ((x101 - xf) * ((y111 - yf) * (QL(m, i, j, k + 1) + QL(m, i, j + 1, k + 1))) +
                                       (xf - x001) * ((yf - y101) * (QL(m, i + 1, j, k + 1) + QL(m, i + 1, j + 1, k + 1))));
                    uvw[m - 2] = down + (up - down) / (z001 - z000) * (zf - z000);
                }
                const double u_x = uvw[0] * vel_scale, u_y = uvw[1] * vel_scale, u_z = uvw[2] * vel_scale;
                const double u_theta = omega * radius + u_y * sin(angle_aux) - u_z * cos(angle_aux);
                const double phi = atan2(u_x, u_theta);
                const double aoa = phi * 180.0 / pi - twist;
                double cl_b, cd_b, cl_a, cd_a;
                polar_lookup(&o->pol[T->polar_before[jk]], aoa, &cl_b, &cd_b);
                polar_lookup(&o->pol[T->polar_after[jk]], aoa, &cl_a, &cd_a);
                const double cl = (cl_a - cl_b) * inter + cl_b;
                const double cd = (cd_a - cd_b) * inter + cd_b;
                const double urel = sqrt(u_theta * u_theta + u_x * u_x);
                const double fL = 0.5 * cl * rho * (urel * urel) * chord;
                const double fD = 0.5 * cd * rho * (urel * urel) * chord;
                double* f = fl + 6 * ((long)(bld - 1) * ne + (elem - 1));
                f[0] = xf;
                f[1] = yf;
                f[2] = zf;
                f[3] = fL * cos(phi) + fD * sin(phi);
                f[4] = fL * sin(angle_aux) * sin(phi) - fD * sin(angle_aux) * sin(phi);
                f[5] = fL * cos(angle_aux) * sin(phi) - fD * cos(angle_aux) * cos(phi);
            }
        }
        /* Gaussian spreading onto the kernel nodes (:355-470) */
        const double deltar = fabs(T->tur_rad - T->hub_rad) / T->nelems;
        for (int c = 0; c < T->n_ker; c++) {
            const int ik = T->ker_nodes[3 * c], jk2 = T->ker_nodes[3 * c + 1], kk = T->ker_nodes[3 * c + 2];
            const double xk = T->ker_pos[3 * c], yk = T->ker_pos[3 * c + 1], zk = T->ker_pos[3 * c + 2];
            const double radius = T->ker_dat[3 * c] * L;
            const double dV = T->ker_dat[3 * c + 2] * (L * L * L);
            if (!(ik >= d->is && ik <= d->ie && jk2 >= d->js && jk2 <= d->je && kk >= d->ks && kk <= d->ke)) continue;
            if (!(radius > T->hub_rad)) continue;
            double fxa = 0.0, fya = 0.0, fza = 0.0;
            for (int bld = 1; bld <= T->nblades; bld++)
                for (int elem = 1; elem <= ne; elem++) {
                    const double* f = fl + 6 * ((long)(bld - 1) * ne + (elem - 1));
                    const double rdis = sqrt((xk - f[0]) * (xk - f[0]) + (yk - f[1]) * (yk - f[1]) +
                                             (zk - f[2]) * (zk - f[2])) * L;
                    const double eps = 2 * pow(dV, 1.0 / 3.0);
                    const double ker = 1 / ((eps * eps * eps) * pow(pi, 3.0 / 2.0)) * exp(-(rdis * rdis) / (eps * eps));
                    fxa = fxa + ker * f[3] * deltar;
                    fya = fya + ker * f[4] * deltar;
                    fza = fza + ker * f[5] * deltar;
                }
            const double s = L / (rho * (vel_scale * vel_scale));
            const double ajv = d->aj[IX(ik, jk2, kk)];
            d->rh[4 * IX(ik, jk2, kk) + 1] = d->rh[4 * IX(ik, jk2, kk) + 1] + fxa * s * (1.0 / ajv);
            d->rh[4 * IX(ik, jk2, kk) + 2] = d->rh[4 * IX(ik, jk2, kk) + 2] + fya * s * (1.0 / ajv);
            d->rh[4 * IX(ik, jk2, kk) + 3] = d->rh[4 * IX(ik, jk2, kk) + 3] + fza * s * (1.0 / ajv);
        }
        free(fl);
    }
#undef QL
}

static int solver_daf(or_state* o, int n, int recalc, int calc_pk) {
    daf_t dd, *d = &dd;
    LEVEL(n);
    long N = (long)ni * nj * nk;
    int i, j, k, m;
    if (ni < 5 || nj < 5 || nk < 5) return -1;
    memset(d, 0, sizeof(*d));
    d->o = o; d->n = n; d->ni = ni; d->nj = nj; d->nk = nk;
    d->dc = o->dc[n]; d->de = o->de[n]; d->dz = o->dz[n];
    d->is = 2; d->ie = ni - 1; d->js = 2; d->je = nj - 1; d->ks = 2; d->ke = nk - 1;
    d->idend = d->ie;
    d->q = o->q + 5 * off; d->qn = o->qn + 5 * off; d->qnm1 = o->qnm1 + 5 * off;
    d->rh = o->rh + 4 * off; d->pk = o->pk + 4 * off;
    d->csi = o->csi + 3 * off; d->eta = o->eta + 3 * off; d->zet = o->zet + 3 * off;
    d->aj = o->aj + off; d->xnut = o->xnut + off;
    d->qold_af = zalloc(4 * N); d->diss = zalloc(4 * N); d->visc = zalloc(3 * N);
    d->ucn = zalloc(3 * N); d->fv = zalloc(3 * N); d->dtau = zalloc(N);
    d->mai = zalloc(16 * N); d->n1i = zalloc(16 * N); d->n2i = zalloc(16 * N); d->mc = zalloc(16 * N);
    d->spr = zalloc(3 * N);
    memset(d->rh, 0, 4 * N * sizeof(double));
    for (k = 1; k <= nk; k++)
        for (j = 1; j <= nj; j++)
            for (i = 1; i <= ni; i++)
                for (m = 1; m <= 4; m++) d->qold_af[4 * IX(i, j, k) + m - 1] = Q(m, i, j, k);
    rhs_contra_j(d);
    rhs_daf_dtau(d);
    rhs_viscous(d);
    rhs_diss_p(d);
    rhs_flux_sans_convec(d);
    rhs_convec_quick(d, n == 1 ? 2 : 1);
    rhs_unst_visc_diss(d);
    if (n > 1) rhs_pk(d, calc_pk == 1 ? 1 : 0);
    zero_rh_boundaries(d);
    /* characteristic outlet on the i faces of the fine level (:322-365) */
    if (n == 1 && o->p.btype[0] == 5) {
        mg_brhs_plane(d, 1);
        nonreflect_plane(d, 1, 1);
    }
    if (n == 1 && o->p.btype[1] == 5) {
        mg_brhs_plane(d, ni);
        nonreflect_plane(d, ni, 2);
    }
    /* the outlet plane joins the sweeps and the update (:370-373) */
    if (n == 1 && o->p.btype[1] == 5) d->idend = d->ie + 1;
    rhs_lines(d);
    rhs_modal_matrices(d);
    rhs_diag_solver(d);
    for (k = d->ks; k <= d->ke; k++)
        for (j = d->js; j <= d->je; j++)
            for (i = d->is; i <= d->idend; i++)
                for (m = 1; m <= 4; m++) Q(m, i, j, k) = d->qold_af[4 * IX(i, j, k) + m - 1] + RH(m, i, j, k);
    if (n == 1) bcond_fm(o, d->q);
    if (recalc == 1) {
        /* recompute rh at the new q; dtau and diss(1) are NOT recomputed (:432-484) */
        rhs_contra_j(d);
        rhs_viscous(d);
        rhs_flux_sans_convec(d);
        rhs_convec_quick(d, n > 1 ? 1 : 2);
        rhs_unst_visc_diss(d);
        if (n > 1) rhs_pk(d, 0);
        zero_rh_boundaries(d);
    }
    free(d->qold_af); free(d->diss); free(d->visc); free(d->ucn); free(d->fv); free(d->dtau);
    free(d->mai); free(d->n1i); free(d->n2i); free(d->mc); free(d->spr);
    return 0;
}

/* residual assembly only (solver_daf.F90:195-315): rh of level n from the
 * stored state, nothing else changed (used by the closed-form checks) */
/* the local pseudo time step of level n (rhs_daf_dtau, or met_daf_dtau's
 * viscous limits per p.dtau_formula) into out (the level's node count) */
int or_dtau(or_state* o, int n, double* out) {
    daf_t dd, *d = &dd;
    if (n < 1 || n > o->ng) return -1;
    {
        LEVEL(n);
        memset(d, 0, sizeof(*d));
        d->o = o; d->n = n; d->ni = ni; d->nj = nj; d->nk = nk;
        d->dc = o->dc[n]; d->de = o->de[n]; d->dz = o->dz[n];
        d->csi = o->csi + 3 * off; d->eta = o->eta + 3 * off; d->zet = o->zet + 3 * off;
        d->aj = o->aj + off; d->xnut = o->xnut + off;
        d->dtau = out;
        rhs_daf_dtau(d);
    }
    return 0;
}

int or_residual(or_state* o, int n, int calc_pk) {
    daf_t dd, *d = &dd;
    if (n < 1 || n > o->ng) return -1;
    {
        LEVEL(n);
        long N = (long)ni * nj * nk;
        if (ni < 5 || nj < 5 || nk < 5) return -1;
        memset(d, 0, sizeof(*d));
        d->o = o; d->n = n; d->ni = ni; d->nj = nj; d->nk = nk;
        d->dc = o->dc[n]; d->de = o->de[n]; d->dz = o->dz[n];
        d->is = 2; d->ie = ni - 1; d->js = 2; d->je = nj - 1; d->ks = 2; d->ke = nk - 1;
        d->idend = d->ie;
        d->q = o->q + 5 * off; d->qn = o->qn + 5 * off; d->qnm1 = o->qnm1 + 5 * off;
        d->rh = o->rh + 4 * off; d->pk = o->pk + 4 * off;
        d->csi = o->csi + 3 * off; d->eta = o->eta + 3 * off; d->zet = o->zet + 3 * off;
        d->aj = o->aj + off; d->xnut = o->xnut + off;
        d->diss = zalloc(4 * N); d->visc = zalloc(3 * N); d->ucn = zalloc(3 * N); d->fv = zalloc(3 * N);
        d->dtau = zalloc(N);
        memset(d->rh, 0, 4 * N * sizeof(double));
        rhs_contra_j(d);
        rhs_daf_dtau(d);
        rhs_viscous(d);
        rhs_diss_p(d);
        rhs_flux_sans_convec(d);
        rhs_convec_quick(d, n == 1 ? 2 : 1);
        rhs_unst_visc_diss(d);
        if (n > 1) rhs_pk(d, calc_pk == 1 ? 1 : 0);
        zero_rh_boundaries(d);
        free(d->diss); free(d->visc); free(d->ucn); free(d->fv); free(d->dtau);
    }
    return 0;
}

int or_solver_daf(or_state* o, int n, int decide_recalc_rh, int decide_calc_pk) {
    if (n < 1 || n > o->ng) return -1;
    return solver_daf(o, n, decide_recalc_rh, decide_calc_pk);
}

#undef Q
#undef QN
#undef QNM1
#undef RH
#undef PK
#undef UCN
#undef FV
#undef VISC
#undef DISS
#undef DTAU
#undef SPR
#undef MAT

/* ------------------------------------------------------------------ */
/* des_eddy and the rhs_sa_* procedures (src/des_eddy.F90)              */
/* ------------------------------------------------------------------ */
/* Spalart-Allmaras constants (des_eddy.F90:63-73) */
static const double cb1 = 0.1355, cb2 = 0.622, sig = 2.0 / 3.0, capa = 0.41, cw1 = 3.23906782, cw2 = 0.3,
                    cw3h = 64.0, cv1c = 357.911, cv2 = 5.0;

typedef struct {
    const or_state* o;
    int ni, nj, nk, is, ie, js, je, ks, ke;
    double dc, de, dz;
    double* q; /* stride 5, full */
    const double *qn, *qnm1; /* stride 5, component 5 used */
    const double *csi, *eta, *zet, *aj, *wd, *dtev;
    double* xnut;
    double *ucn, *qold, *visc, *dnu, *psa, *dk, *wo, *fw, *fv, *rh;
} sa_t;

#define Q(m, i, j, k) d->q[5 * IX(i, j, k) + (m)-1]
#define Q5(i, j, k) d->q[5 * IX(i, j, k) + 4]
#define UCN(c, i, j, k) d->ucn[3 * IX(i, j, k) + (c)-1]
#define S1(a, i, j, k) d->a[IX(i, j, k)]
#define SLV const int ni = d->ni, nj = d->nj, nk = d->nk

static void sa_contra_j(sa_t* d) {
    SLV;
    int i, j, k;
    for (k = 1; k <= nk; k++)
        for (j = 1; j <= nj; j++)
            for (i = 1; i <= ni; i++) {
                double tmp = 1.0 / AJ(i, j, k);
                UCN(1, i, j, k) = tmp * (CSI(1, i, j, k) * Q(2, i, j, k) + CSI(2, i, j, k) * Q(3, i, j, k) + CSI(3, i, j, k) * Q(4, i, j, k));
                UCN(2, i, j, k) = tmp * (ETA(1, i, j, k) * Q(2, i, j, k) + ETA(2, i, j, k) * Q(3, i, j, k) + ETA(3, i, j, k) * Q(4, i, j, k));
                UCN(3, i, j, k) = tmp * (ZET(1, i, j, k) * Q(2, i, j, k) + ZET(2, i, j, k) * Q(3, i, j, k) + ZET(3, i, j, k) * Q(4, i, j, k));
            }
}

/* rhs_sa_source (src/rhs_sa_source.F90:2-95): interior only, early return */
static void rhs_sa_source(sa_t* d) {
    SLV;
    int i, j, k;
    const double dc2 = pt5 * d->dc, de2 = pt5 * d->de, dz2 = pt5 * d->dz;
    (void)nk;
    for (k = d->ks; k <= d->ke; k++)
        for (j = d->js; j <= d->je; j++)
            for (i = d->is; i <= d->ie; i++) {
                double uc = dc2 * (Q(2, i + 1, j, k) - Q(2, i - 1, j, k));
                double vc = dc2 * (Q(3, i + 1, j, k) - Q(3, i - 1, j, k));
                double wc = dc2 * (Q(4, i + 1, j, k) - Q(4, i - 1, j, k));
                double ue = de2 * (Q(2, i, j + 1, k) - Q(2, i, j - 1, k));
                double ve = de2 * (Q(3, i, j + 1, k) - Q(3, i, j - 1, k));
                double we = de2 * (Q(4, i, j + 1, k) - Q(4, i, j - 1, k));
                double uz = dz2 * (Q(2, i, j, k + 1) - Q(2, i, j, k - 1));
                double vz = dz2 * (Q(3, i, j, k + 1) - Q(3, i, j, k - 1));
                double wz = dz2 * (Q(4, i, j, k + 1) - Q(4, i, j, k - 1));
                double u2 = uc * CSI(2, i, j, k) + ue * ETA(2, i, j, k) + uz * ZET(2, i, j, k);
                double u3 = uc * CSI(3, i, j, k) + ue * ETA(3, i, j, k) + uz * ZET(3, i, j, k);
                double v1 = vc * CSI(1, i, j, k) + ve * ETA(1, i, j, k) + vz * ZET(1, i, j, k);
                double v3 = vc * CSI(3, i, j, k) + ve * ETA(3, i, j, k) + vz * ZET(3, i, j, k);
                double w1 = wc * CSI(1, i, j, k) + we * ETA(1, i, j, k) + wz * ZET(1, i, j, k);
                double w2 = wc * CSI(2, i, j, k) + we * ETA(2, i, j, k) + wz * ZET(2, i, j, k);
                double w12 = pt5 * (u2 - v1), w13 = pt5 * (u3 - w1), w23 = pt5 * (v3 - w2);
                S1(wo, i, j, k) = 2.0 * (w12 * w12 + w13 * w13 + w23 * w23);
            }
}

/* rhs_sa_grdnu (src/rhs_sa_grdnu.F90:1-63) */
static void rhs_sa_grdnu(sa_t* d) {
    SLV;
    int i, j, k;
    const double dc2 = pt5 * d->dc, de2 = pt5 * d->de, dz2 = pt5 * d->dz;
    long N = (long)ni * nj * nk;
    double* q5j = zalloc(N);
#define Q5J(i, j, k) q5j[IX(i, j, k)]
    for (k = 1; k <= nk; k++)
        for (j = 1; j <= nj; j++)
            for (i = 1; i <= ni; i++) Q5J(i, j, k) = Q5(i, j, k) / AJ(i, j, k);
    for (k = d->ks; k <= d->ke; k++)
        for (j = d->js; j <= d->je; j++)
            for (i = d->is; i <= d->ie; i++) {
                double dnu1 = dc2 * (Q5J(i + 1, j, k) * CSI(1, i + 1, j, k) - Q5J(i - 1, j, k) * CSI(1, i - 1, j, k));
                double dnu2 = dc2 * (Q5J(i + 1, j, k) * CSI(2, i + 1, j, k) - Q5J(i - 1, j, k) * CSI(2, i - 1, j, k));
                double dnu3 = dc2 * (Q5J(i + 1, j, k) * CSI(3, i + 1, j, k) - Q5J(i - 1, j, k) * CSI(3, i - 1, j, k));
                double aj2;
                dnu1 = dnu1 + de2 * (Q5J(i, j + 1, k) * ETA(1, i, j + 1, k) - Q5J(i, j - 1, k) * ETA(1, i, j - 1, k));
                dnu2 = dnu2 + de2 * (Q5J(i, j + 1, k) * ETA(2, i, j + 1, k) - Q5J(i, j - 1, k) * ETA(2, i, j - 1, k));
                dnu3 = dnu3 + de2 * (Q5J(i, j + 1, k) * ETA(3, i, j + 1, k) - Q5J(i, j - 1, k) * ETA(3, i, j - 1, k));
                dnu1 = dnu1 + dz2 * (Q5J(i, j, k + 1) * ZET(1, i, j, k + 1) - Q5J(i, j, k - 1) * ZET(1, i, j, k - 1));
                dnu2 = dnu2 + dz2 * (Q5J(i, j, k + 1) * ZET(2, i, j, k + 1) - Q5J(i, j, k - 1) * ZET(2, i, j, k - 1));
                dnu3 = dnu3 + dz2 * (Q5J(i, j, k + 1) * ZET(3, i, j, k + 1) - Q5J(i, j, k - 1) * ZET(3, i, j, k - 1));
                aj2 = AJ(i, j, k) * AJ(i, j, k);
                S1(dnu, i, j, k) = aj2 * (dnu1 * dnu1 + dnu2 * dnu2 + dnu3 * dnu3);
            }
#undef Q5J
    free(q5j);
}

/* rhs_sa_prod (src/rhs_sa_prod.F90:1-105) */
static void rhs_sa_prod(sa_t* d) {
    SLV;
    int i, j, k;
    const double rei = 1.0 / d->o->p.ren;
    const double vtc_col = 0.00001, one_sixth = 1.0 / 6.0;
    (void)nk;
    for (k = d->ks; k <= d->ke; k++)
        for (j = d->js; j <= d->je; j++)
            for (i = d->is; i <= d->ie; i++) {
                if (S1(wd, i, j, k) == 0.0) {
                    S1(psa, i, j, k) = 0.0;
                } else {
                    double q5 = Q5(i, j, k);
                    double chi = q5 / rei;
                    double chic = chi * chi * chi;
                    double fv1 = chic / (chic + cv1c);
                    double fv2 = pow(1.0 + chi / cv2, -3.0);
                    double fv3 = ((1.0 + chi * fv1) * (1.0 - fv2)) / chi;
                    double vtcm = sqrt(2.0 * S1(wo, i, j, k));
                    double cawd = capa * capa * S1(wd, i, j, k) * S1(wd, i, j, k);
                    double vtc_org = fv3 * vtcm + q5 * fv2 / cawd;
                    double vtc = dmax2(vtc_org, vtc_col);
                    double rr, rr6, grr_org, grr, grr6, vwd;
                    S1(psa, i, j, k) = cb1 * vtc * q5;
                    rr = q5 / (vtc * cawd);
                    rr6 = rr * rr * rr * rr * rr * rr;
                    grr_org = rr + cw2 * (rr6 - rr);
                    grr = dmax2(0.0, grr_org);
                    grr6 = grr * grr * grr * grr * grr * grr;
                    S1(fw, i, j, k) = grr * pow((1.0 + cw3h) / (grr6 + cw3h), one_sixth);
                    vwd = q5 / S1(wd, i, j, k);
                    S1(psa, i, j, k) = S1(psa, i, j, k) - cw1 * S1(fw, i, j, k) * vwd * vwd;
                    S1(psa, i, j, k) = S1(psa, i, j, k) + cb2 * S1(dnu, i, j, k) / sig;
                }
            }
}

/* rhs_sa_viscous (src/rhs_sa_viscous.F90:2-192) */
static void rhs_sa_viscous(sa_t* d) {
    SLV;
    int i, j, k;
    const double dc = d->dc, de = d->de, dz = d->dz;
    const double dc4 = pt25 * dc, de4 = pt25 * de, dz4 = pt25 * dz;
    const double rei = 1.0 / d->o->p.ren;
    long N = (long)ni * nj * nk;
    double* ret = zalloc(N);
#define RET(i, j, k) ret[IX(i, j, k)]
    for (k = 1; k <= nk; k++)
        for (j = 1; j <= nj; j++)
            for (i = 1; i <= ni; i++) RET(i, j, k) = (rei + Q5(i, j, k)) / sig;
    for (k = d->ks; k <= d->ke; k++)
        for (j = d->js; j <= d->je; j++)
            for (i = d->is - 1; i <= d->ie; i++) {
                double vcs = dc * (Q5(i + 1, j, k) - Q5(i, j, k));
                double vet = de4 * (Q5(i + 1, j + 1, k) - Q5(i + 1, j - 1, k) + Q5(i, j + 1, k) - Q5(i, j - 1, k));
                double vze = dz4 * (Q5(i + 1, j, k + 1) - Q5(i + 1, j, k - 1) + Q5(i, j, k + 1) - Q5(i, j, k - 1));
                double a1 = CSI(1, i + 1, j, k) + CSI(1, i, j, k), a2 = CSI(2, i + 1, j, k) + CSI(2, i, j, k), a3 = CSI(3, i + 1, j, k) + CSI(3, i, j, k);
                double b1 = ETA(1, i + 1, j, k) + ETA(1, i, j, k), b2 = ETA(2, i + 1, j, k) + ETA(2, i, j, k), b3 = ETA(3, i + 1, j, k) + ETA(3, i, j, k);
                double c1 = ZET(1, i + 1, j, k) + ZET(1, i, j, k), c2 = ZET(2, i + 1, j, k) + ZET(2, i, j, k), c3 = ZET(3, i + 1, j, k) + ZET(3, i, j, k);
                double cc1 = a1 * a1 + a2 * a2 + a3 * a3, cc2 = a1 * b1 + a2 * b2 + a3 * b3, cc3 = a1 * c1 + a2 * c2 + a3 * c3;
                double vtj = pt25 * (RET(i + 1, j, k) + RET(i, j, k)) / (AJ(i + 1, j, k) + AJ(i, j, k));
                S1(fv, i, j, k) = vtj * (cc1 * vcs + cc2 * vet + cc3 * vze);
            }
    for (k = d->ks; k <= d->ke; k++)
        for (j = d->js; j <= d->je; j++)
            for (i = d->is; i <= d->ie; i++) S1(visc, i, j, k) = S1(fv, i, j, k) - S1(fv, i - 1, j, k);
    for (k = d->ks; k <= d->ke; k++)
        for (j = d->js - 1; j <= d->je; j++)
            for (i = d->is; i <= d->ie; i++) {
                double vcs = dc4 * (Q5(i + 1, j + 1, k) - Q5(i - 1, j + 1, k) + Q5(i + 1, j, k) - Q5(i - 1, j, k));
                double vet = de * (Q5(i, j + 1, k) - Q5(i, j, k));
                double vze = dz4 * (Q5(i, j + 1, k + 1) - Q5(i, j + 1, k - 1) + Q5(i, j, k + 1) - Q5(i, j, k - 1));
                double a1 = CSI(1, i, j + 1, k) + CSI(1, i, j, k), a2 = CSI(2, i, j + 1, k) + CSI(2, i, j, k), a3 = CSI(3, i, j + 1, k) + CSI(3, i, j, k);
                double b1 = ETA(1, i, j + 1, k) + ETA(1, i, j, k), b2 = ETA(2, i, j + 1, k) + ETA(2, i, j, k), b3 = ETA(3, i, j + 1, k) + ETA(3, i, j, k);
                double c1 = ZET(1, i, j + 1, k) + ZET(1, i, j, k), c2 = ZET(2, i, j + 1, k) + ZET(2, i, j, k), c3 = ZET(3, i, j + 1, k) + ZET(3, i, j, k);
                double cc1 = b1 * a1 + b2 * a2 + b3 * a3, cc2 = b1 * b1 + b2 * b2 + b3 * b3, cc3 = b1 * c1 + b2 * c2 + b3 * c3;
                double vtj = pt25 * (RET(i, j + 1, k) + RET(i, j, k)) / (AJ(i, j + 1, k) + AJ(i, j, k));
                S1(fv, i, j, k) = vtj * (cc1 * vcs + cc2 * vet + cc3 * vze);
            }
    for (k = d->ks; k <= d->ke; k++)
        for (j = d->js; j <= d->je; j++)
            for (i = d->is; i <= d->ie; i++) S1(visc, i, j, k) = S1(visc, i, j, k) + S1(fv, i, j, k) - S1(fv, i, j - 1, k);
    for (k = d->ks - 1; k <= d->ke; k++)
        for (j = d->js; j <= d->je; j++)
            for (i = d->is; i <= d->ie; i++) {
                double vcs = dc4 * (Q5(i + 1, j, k + 1) - Q5(i - 1, j, k + 1) + Q5(i + 1, j, k) - Q5(i - 1, j, k));
                double vet = de4 * (Q5(i, j + 1, k + 1) - Q5(i, j - 1, k + 1) + Q5(i, j + 1, k) - Q5(i, j - 1, k));
                double vze = dz * (Q5(i, j, k + 1) - Q5(i, j, k));
                double a1 = CSI(1, i, j, k + 1) + CSI(1, i, j, k), a2 = CSI(2, i, j, k + 1) + CSI(2, i, j, k), a3 = CSI(3, i, j, k + 1) + CSI(3, i, j, k);
                double b1 = ETA(1, i, j, k + 1) + ETA(1, i, j, k), b2 = ETA(2, i, j, k + 1) + ETA(2, i, j, k), b3 = ETA(3, i, j, k + 1) + ETA(3, i, j, k);
                double c1 = ZET(1, i, j, k + 1) + ZET(1, i, j, k), c2 = ZET(2, i, j, k + 1) + ZET(2, i, j, k), c3 = ZET(3, i, j, k + 1) + ZET(3, i, j, k);
                double cc1 = c1 * a1 + c2 * a2 + c3 * a3, cc2 = c1 * b1 + c2 * b2 + c3 * b3, cc3 = c1 * c1 + c2 * c2 + c3 * c3;
                double vtj = pt25 * (RET(i, j, k + 1) + RET(i, j, k)) / (AJ(i, j, k + 1) + AJ(i, j, k));
                S1(fv, i, j, k) = vtj * (cc1 * vcs + cc2 * vet + cc3 * vze);
            }
    for (k = d->ks; k <= d->ke; k++)
        for (j = d->js; j <= d->je; j++)
            for (i = d->is; i <= d->ie; i++) S1(visc, i, j, k) = S1(visc, i, j, k) + S1(fv, i, j, k) - S1(fv, i, j, k - 1);
#undef RET
    free(ret);
}

/* rhs_sa_flux (src/rhs_sa_flux.F90:2-112): first-order upwind FDS */
static void rhs_sa_flux(sa_t* d) {
    SLV;
    int i, j, k, dir;
    (void)nk;
    for (dir = 1; dir <= 3; dir++) {
        const int di = dir == 1, dj = dir == 2, dk = dir == 3;
        for (k = d->ks - dk; k <= d->ke; k++)
            for (j = d->js - dj; j <= d->je; j++)
                for (i = d->is - di; i <= d->ie; i++) {
                    double u1 = UCN(dir, i + di, j + dj, k + dk), u0 = UCN(dir, i, j, k);
                    double q1 = Q5(i + di, j + dj, k + dk), q0 = Q5(i, j, k);
                    double eigh = pt5 * fabs(u1 + u0);
                    S1(fv, i, j, k) = pt5 * (q1 * u1 + q0 * u0 - eigh * (q1 - q0));
                }
        for (k = d->ks; k <= d->ke; k++)
            for (j = d->js; j <= d->je; j++)
                for (i = d->is; i <= d->ie; i++) {
                    if (dir == 1)
                        S1(rh, i, j, k) = S1(fv, i, j, k) - S1(fv, i - 1, j, k);
                    else
                        S1(rh, i, j, k) = S1(rh, i, j, k) + S1(fv, i, j, k) - S1(fv, i - di, j - dj, k - dk);
                }
    }
}

/* rhs_sa_unst_visc (src/rhs_sa_unst_visc.F90:1-32) */
static void rhs_sa_unst_visc(sa_t* d) {
    SLV;
    int i, j, k;
    const double es = d->o->e_source, delti = d->o->p.delti;
    (void)nk;
    for (k = d->ks; k <= d->ke; k++)
        for (j = d->js; j <= d->je; j++)
            for (i = d->is; i <= d->ie; i++) {
                double dtj = 2.0 * delti * AJ(i, j, k);
                double dq5dt = es * (3.0 * Q5(i, j, k) - 4.0 * d->qn[5 * IX(i, j, k) + 4] + d->qnm1[5 * IX(i, j, k) + 4]) / dtj;
                S1(rh, i, j, k) = S1(rh, i, j, k) + dq5dt - S1(visc, i, j, k) - S1(psa, i, j, k) / AJ(i, j, k);
            }
}

/* rhs_sa_implicit_source (src/rhs_sa_implicit_source.F90:2-55): all nodes */
static void rhs_sa_implicit_source(sa_t* d) {
    SLV;
    int i, j, k;
    for (k = 1; k <= nk; k++)
        for (j = 1; j <= nj; j++)
            for (i = 1; i <= ni; i++) {
                S1(dk, i, j, k) = 1.0;
                S1(rh, i, j, k) = -S1(dtev, i, j, k) * AJ(i, j, k) * S1(rh, i, j, k);
            }
}

/* rhs_sa_adi_solver (src/rhs_sa_adi_solver.F90:2-248) */
static void rhs_sa_adi_solver(sa_t* d) {
    SLV;
    int i, j, k, s, dir;
    const double rei = 1.0 / d->o->p.ren;
    int maxlen = ni > nj ? ni : nj;
    double *aw, *ap, *ae, *aq, *av;
    if (nk > maxlen) maxlen = nk;
    aw = zalloc(maxlen); ap = zalloc(maxlen); ae = zalloc(maxlen); aq = zalloc(maxlen); av = zalloc(maxlen + 1);
    for (dir = 1; dir <= 3; dir++) {
        const int di = dir == 1, dj = dir == 2, dk = dir == 3;
        const double* met = dir == 1 ? d->csi : dir == 2 ? d->eta : d->zet;
        int a0, a1, b0, b1, len, s0;
        /* lines: dir 1 over (j,k); dir 2 over (i,k); dir 3 over (i,j) */
        if (dir == 1) { a0 = d->js; a1 = d->je; b0 = d->ks; b1 = d->ke; s0 = d->is; len = d->ie - d->is + 1; }
        else if (dir == 2) { a0 = d->is; a1 = d->ie; b0 = d->ks; b1 = d->ke; s0 = d->js; len = d->je - d->js + 1; }
        else { a0 = d->is; a1 = d->ie; b0 = d->js; b1 = d->je; s0 = d->ks; len = d->ke - d->ks + 1; }
        {
            int la, lb;
            for (lb = b0; lb <= b1; lb++)
                for (la = a0; la <= a1; la++) {
                    /* av(s) for faces s+1/2, s = s0-1 .. s0+len-1 -> av[s - s0 + 1] */
                    for (s = s0 - 1; s <= s0 + len - 1; s++) {
                        int i0, j0, k0;
                        double gx, gy, gz, ev, ucna;
                        long l0, l1;
                        if (dir == 1) { i0 = s; j0 = la; k0 = lb; }
                        else if (dir == 2) { i0 = la; j0 = s; k0 = lb; }
                        else { i0 = la; j0 = lb; k0 = s; }
                        l0 = IX(i0, j0, k0);
                        l1 = IX(i0 + di, j0 + dj, k0 + dk);
                        gx = met[3 * l1 + 0] + met[3 * l0 + 0];
                        gy = met[3 * l1 + 1] + met[3 * l0 + 1];
                        gz = met[3 * l1 + 2] + met[3 * l0 + 2];
                        ev = (rei + pt5 * (d->q[5 * l1 + 4] + d->q[5 * l0 + 4])) / sig;
                        ucna = pt5 * fabs(d->ucn[3 * l1 + dir - 1] + d->ucn[3 * l0 + dir - 1]);
                        av[s - s0 + 1] = ucna + ev * (gx * gx + gy * gy + gz * gz) / (d->aj[l1] + d->aj[l0]);
                    }
                    for (s = s0; s <= s0 + len - 1; s++) {
                        int i0, j0, k0;
                        long l, lm, lp;
                        double dtj;
                        if (dir == 1) { i0 = s; j0 = la; k0 = lb; }
                        else if (dir == 2) { i0 = la; j0 = s; k0 = lb; }
                        else { i0 = la; j0 = lb; k0 = s; }
                        l = IX(i0, j0, k0);
                        lm = IX(i0 - di, j0 - dj, k0 - dk);
                        lp = IX(i0 + di, j0 + dj, k0 + dk);
                        dtj = pt5 * d->dtev[l] * d->aj[l] / d->dk[l];
                        aw[s - s0] = -dtj * (d->ucn[3 * lm + dir - 1] + av[s - s0]);
                        ap[s - s0] = 1.0 + dtj * (+av[s - s0] + av[s - s0 + 1]);
                        ae[s - s0] = dtj * (d->ucn[3 * lp + dir - 1] - av[s - s0 + 1]);
                        aq[s - s0] = dir == 1 ? d->rh[l] / d->dk[l] : d->rh[l];
                    }
                    or_gtsv(len, 1, aw + 1, ap, ae, aq, len);
                    for (s = s0; s <= s0 + len - 1; s++) {
                        int i0, j0, k0;
                        if (dir == 1) { i0 = s; j0 = la; k0 = lb; }
                        else if (dir == 2) { i0 = la; j0 = s; k0 = lb; }
                        else { i0 = la; j0 = lb; k0 = s; }
                        d->rh[IX(i0, j0, k0)] = aq[s - s0];
                    }
                }
        }
    }
    free(aw); free(ap); free(ae); free(aq); free(av);
}

/* rhs_sa_bcond (src/rhs_sa_bcond.F90:2-294) */
#define SRAT(b, a, c, na) o->rat[b][((a)-1) + (long)(na) * ((c)-1)]
static void rhs_sa_bcond(sa_t* d) {
    SLV;
    const or_state* o = d->o;
    int i, j, k;
    const double rei = 1.0 / o->p.ren;
    const int* bt = o->p.btype;
    /* i = 1 */
    switch (bt[0]) {
        case 1: for (k = 1; k <= nk; k++) for (j = 1; j <= nj; j++) Q5(1, j, k) = 0.0; break;
        case 2: case 3:
            for (k = 1; k <= nk; k++) for (j = 1; j <= nj; j++) Q5(1, j, k) = o->sa[0][0] * Q5(2, j, k) + o->sb[0][0] * Q5(3, j, k);
            break;
        case 6: case 7: for (k = 1; k <= nk; k++) for (j = 1; j <= nj; j++) Q5(1, j, k) = Q5(d->ie, j, k); break;
        case 5: /* characteristic outlet: extrapolated on the grid ratio */
            for (k = 1; k <= nk; k++) for (j = 1; j <= nj; j++) Q5(1, j, k) = (1.0 + SRAT(0, j, k, nj)) * Q5(2, j, k) - SRAT(0, j, k, nj) * Q5(3, j, k);
            break;
        default: break;
    }
    switch (bt[1]) {
        case 1: for (k = 1; k <= nk; k++) for (j = 1; j <= nj; j++) Q5(ni, j, k) = 0.0; break;
        case 2: case 3:
            for (k = 1; k <= nk; k++) for (j = 1; j <= nj; j++) Q5(ni, j, k) = o->sa[1][0] * Q5(ni - 1, j, k) + o->sb[1][0] * Q5(ni - 2, j, k);
            break;
        case 6: case 7: for (k = 1; k <= nk; k++) for (j = 1; j <= nj; j++) Q5(ni, j, k) = Q5(d->is, j, k); break;
        case 8: for (k = 1; k <= nk; k++) for (j = 1; j <= nj; j++) Q5(ni, j, k) = Q5(ni - 1, j, k); break;
        case 5:
            for (k = 1; k <= nk; k++) for (j = 1; j <= nj; j++) Q5(ni, j, k) = (1.0 + SRAT(1, j, k, nj)) * Q5(ni - 1, j, k) - SRAT(1, j, k, nj) * Q5(ni - 2, j, k);
            break;
        default: break;
    }
    switch (bt[2]) {
        case 1: for (k = 1; k <= nk; k++) for (i = 1; i <= ni; i++) Q5(i, 1, k) = 0.0; break;
        case 2: case 3:
            for (k = 1; k <= nk; k++) for (i = 1; i <= ni; i++) Q5(i, 1, k) = o->sa[2][0] * Q5(i, 2, k) + o->sb[2][0] * Q5(i, 3, k);
            break;
        case 6: case 7: for (k = 1; k <= nk; k++) for (i = 1; i <= ni; i++) Q5(i, 1, k) = Q5(i, d->je, k); break;
        case 5:
            for (k = 1; k <= nk; k++) for (i = 1; i <= ni; i++) Q5(i, 1, k) = (1.0 + SRAT(2, i, k, ni)) * Q5(i, 2, k) - SRAT(2, i, k, ni) * Q5(i, 3, k);
            break;
        default: break;
    }
    switch (bt[3]) {
        case 1: for (k = 1; k <= nk; k++) for (i = 1; i <= ni; i++) Q5(i, nj, k) = 0.0; break;
        case 2: case 3:
            for (k = 1; k <= nk; k++) for (i = 1; i <= ni; i++) Q5(i, nj, k) = o->sa[3][0] * Q5(i, nj - 1, k) + o->sb[3][0] * Q5(i, nj - 2, k);
            break;
        case 6: case 7: for (k = 1; k <= nk; k++) for (i = 1; i <= ni; i++) Q5(i, nj, k) = Q5(i, d->js, k); break;
        case 5:
            for (k = 1; k <= nk; k++) for (i = 1; i <= ni; i++) Q5(i, nj, k) = (1.0 + SRAT(3, i, k, ni)) * Q5(i, nj - 1, k) - SRAT(3, i, k, ni) * Q5(i, nj - 2, k);
            break;
        default: break;
    }
    switch (bt[4]) {
        case 1: for (j = 1; j <= nj; j++) for (i = 1; i <= ni; i++) Q5(i, j, 1) = 0.0; break;
        case 2: case 3:
            for (j = 1; j <= nj; j++) for (i = 1; i <= ni; i++) Q5(i, j, 1) = o->sa[4][0] * Q5(i, j, 2) + o->sb[4][0] * Q5(i, j, 3);
            break;
        case 6: case 7: for (j = 1; j <= nj; j++) for (i = 1; i <= ni; i++) Q5(i, j, 1) = Q5(i, j, d->ke); break;
        case 5:
            for (j = 1; j <= nj; j++) for (i = 1; i <= ni; i++) Q5(i, j, 1) = (1.0 + SRAT(4, i, j, ni)) * Q5(i, j, 2) - SRAT(4, i, j, ni) * Q5(i, j, 3);
            break;
        default: break;
    }
    switch (bt[5]) {
        case 1: for (j = 1; j <= nj; j++) for (i = 1; i <= ni; i++) Q5(i, j, nk) = 0.0; break;
        case 2: case 3:
            for (j = 1; j <= nj; j++) for (i = 1; i <= ni; i++) Q5(i, j, nk) = o->sa[5][0] * Q5(i, j, nk - 1) + o->sb[5][0] * Q5(i, j, nk - 2);
            break;
        case 6: case 7: for (j = 1; j <= nj; j++) for (i = 1; i <= ni; i++) Q5(i, j, nk) = Q5(i, j, d->ks); break;
        case 5:
            for (j = 1; j <= nj; j++) for (i = 1; i <= ni; i++) Q5(i, j, nk) = (1.0 + SRAT(5, i, j, ni)) * Q5(i, j, nk - 1) - SRAT(5, i, j, ni) * Q5(i, j, nk - 2);
            break;
        default: break;
    }
    for (k = 1; k <= nk; k++)
        for (j = 1; j <= nj; j++)
            for (i = 1; i <= ni; i++) {
                double chi, chic, fv1;
                Q5(i, j, k) = dmax2(1.0e-07, Q5(i, j, k));
                chi = Q5(i, j, k) / rei;
                chic = chi * chi * chi;
                fv1 = chic / (chic + cv1c);
                d->xnut[IX(i, j, k)] = Q5(i, j, k) * fv1;
            }
}

static void des_eddy(or_state* o) {
    sa_t dd, *d = &dd;
    LEVEL(1);
    long N = (long)ni * nj * nk;
    int i, j, k, iterk;
    memset(d, 0, sizeof(*d));
    d->o = o; d->ni = ni; d->nj = nj; d->nk = nk;
    d->is = 2; d->ie = ni - 1; d->js = 2; d->je = nj - 1; d->ks = 2; d->ke = nk - 1;
    d->dc = o->dc[1]; d->de = o->de[1]; d->dz = o->dz[1];
    d->q = o->q + 5 * off; d->qn = o->qn + 5 * off; d->qnm1 = o->qnm1 + 5 * off;
    d->csi = o->csi + 3 * off; d->eta = o->eta + 3 * off; d->zet = o->zet + 3 * off;
    d->aj = o->aj + off; d->wd = o->wd + off; d->dtev = o->dtev + off; d->xnut = o->xnut + off;
    d->ucn = zalloc(3 * N); d->qold = zalloc(N); d->visc = zalloc(N); d->dnu = zalloc(N); d->psa = zalloc(N);
    d->dk = zalloc(N); d->wo = zalloc(N); d->fw = zalloc(N); d->fv = zalloc(N); d->rh = zalloc(N);
    for (k = 1; k <= nk; k++)
        for (j = 1; j <= nj; j++)
            for (i = 1; i <= ni; i++) S1(qold, i, j, k) = Q5(i, j, k);
    sa_contra_j(d);
    rhs_sa_source(d);
    for (iterk = 1; iterk <= o->p.itk; iterk++) {
        rhs_sa_grdnu(d);
        rhs_sa_prod(d);
        rhs_sa_viscous(d);
        rhs_sa_flux(d);
        rhs_sa_unst_visc(d);
        rhs_sa_implicit_source(d);
        for (k = 1; k <= nk; k++)
            for (j = 1; j <= nj; j++) S1(rh, 1, j, k) = 0.0, S1(rh, ni, j, k) = 0.0;
        for (k = 1; k <= nk; k++)
            for (i = 1; i <= ni; i++) S1(rh, i, 1, k) = 0.0, S1(rh, i, nj, k) = 0.0;
        for (j = 1; j <= nj; j++)
            for (i = 1; i <= ni; i++) S1(rh, i, j, 1) = 0.0, S1(rh, i, j, nk) = 0.0;
        rhs_sa_adi_solver(d);
        for (k = d->ks; k <= d->ke; k++)
            for (j = d->js; j <= d->je; j++)
                for (i = d->is; i <= d->ie; i++) {
                    double qe5 = S1(qold, i, j, k) + S1(rh, i, j, k);
                    double dum = 1.0 + onept5 * o->e_source * S1(dtev, i, j, k) / o->p.delti;
                    Q5(i, j, k) = Q5(i, j, k) + (-Q5(i, j, k) + qe5) / dum;
                }
        rhs_sa_bcond(d);
        for (k = 1; k <= nk; k++)
            for (j = 1; j <= nj; j++)
                for (i = 1; i <= ni; i++) S1(qold, i, j, k) = Q5(i, j, k);
    }
    free(d->ucn); free(d->qold); free(d->visc); free(d->dnu); free(d->psa);
    free(d->dk); free(d->wo); free(d->fw); free(d->fv); free(d->rh);
}

int or_des_eddy(or_state* o) {
    if (o->img[1] < 5 || o->jmg[1] < 5 || o->kmg[1] < 5) return -1;
    des_eddy(o);
    return 0;
}

#undef Q
#undef Q5
#undef UCN
#undef S1
#undef CSI
#undef ETA
#undef ZET
#undef AJ
#undef XNUT

/* ------------------------------------------------------------------ */
/* multigrid transfers                                                  */
/* ------------------------------------------------------------------ */
static long gidx(const or_state* o, int n, int i, int j, int k) {
    return o->off[n] + (i - 1) + (long)o->img[n] * ((j - 1) + (long)o->jmg[n] * (k - 1));
}

/* mg_zero_pk_boundary (src/mg_driver.F90:407-466) */
static void mg_zero_pk_boundary(or_state* o, int n) {
    LEVEL(n);
    int i, j, k, m;
    for (k = 1; k <= nk; k++)
        for (j = 1; j <= nj; j++)
            for (m = 0; m < 4; m++) o->pk[4 * gidx(o, n, 1, j, k) + m] = 0.0, o->pk[4 * gidx(o, n, ni, j, k) + m] = 0.0;
    for (k = 1; k <= nk; k++)
        for (i = 1; i <= ni; i++)
            for (m = 0; m < 4; m++) o->pk[4 * gidx(o, n, i, 1, k) + m] = 0.0, o->pk[4 * gidx(o, n, i, nj, k) + m] = 0.0;
    for (j = 1; j <= nj; j++)
        for (i = 1; i <= ni; i++)
            for (m = 0; m < 4; m++) o->pk[4 * gidx(o, n, i, j, 1) + m] = 0.0, o->pk[4 * gidx(o, n, i, j, nk) + m] = 0.0;
}

/* mg_nlevel (src/mg_nlevel.F90:2-150).  The 27-point restriction is
 * evaluated at interior coarse nodes only; the reference reads outside the
 * fine level at boundary coarse nodes and those values are zeroed right
 * after by mg_zero_pk_boundary (src/mg_driver.F90:134), so they are set to
 * zero here. */
static void mg_nlevel(or_state* o, int n) {
    LEVEL(n);
    int i, j, k, m;
    const long ic = 1, ie = o->img[n - 1], iz = (long)o->img[n - 1] * o->jmg[n - 1];
    for (k = 1; k <= nk; k++)
        for (j = 1; j <= nj; j++)
            for (i = 1; i <= ni; i++) {
                int kk = (1 << o->p.kcrs) * k - o->p.kcrs, jj = (1 << o->p.jcrs) * j - o->p.jcrs, ii = (1 << o->p.icrs) * i - o->p.icrs;
                long ll = gidx(o, n - 1, ii, jj, kk), l = gidx(o, n, i, j, k);
                for (m = 0; m < 4; m++) o->q[5 * l + m] = o->q[5 * ll + m];
                o->xnut[l] = o->xnut[ll];
            }
    for (k = 1; k <= nk; k++)
        for (j = 1; j <= nj; j++)
            for (i = 1; i <= ni; i++) {
                long l = gidx(o, n, i, j, k);
                for (m = 0; m < 4; m++) o->qold_mg[5 * l + m] = o->q[5 * l + m];
            }
    for (k = 1; k <= nk; k++)
        for (j = 1; j <= nj; j++)
            for (i = 1; i <= ni; i++) {
                int kk = (1 << o->p.kcrs) * k - o->p.kcrs, jj = (1 << o->p.jcrs) * j - o->p.jcrs, ii = (1 << o->p.icrs) * i - o->p.icrs;
                long ll = gidx(o, n - 1, ii, jj, kk), l = gidx(o, n, i, j, k);
                if (i == 1 || i == ni || j == 1 || j == nj || k == 1 || k == nk) {
                    for (m = 0; m < 4; m++) o->pk[4 * l + m] = 0.0;
                    continue;
                }
                for (m = 0; m < 4; m++) {
                    const double* r = o->rh + m;
#define R(off) r[4 * (ll + (off))]
                    o->pk[4 * l + m] =
                        (pt25 * R(ic + ie + iz) + pt5 * R(ic + ie) + pt25 * R(ic + ie - iz) +
                         pt5 * R(ie + iz) + R(ie) + pt5 * R(ie - iz) +
                         pt25 * R(-ic + ie + iz) + pt5 * R(-ic + ie) + pt25 * R(-ic + ie - iz) +
                         pt5 * R(ic + iz) + R(ic) + pt5 * R(ic - iz) +
                         R(iz) + 2.0 * R(0) + R(-iz) +
                         pt5 * R(-ic + iz) + R(-ic) + pt5 * R(-ic - iz) +
                         pt25 * R(ic - ie + iz) + pt5 * R(ic - ie) + pt25 * R(ic - ie - iz) +
                         pt5 * R(-ie + iz) + R(-ie) + pt5 * R(-ie - iz) +
                         pt25 * R(-ic - ie + iz) + pt5 * R(-ic - ie) + pt25 * R(-ic - ie - iz)) *
                        (1.0 / 16.0);
#undef R
                }
            }
}

int or_mg_nlevel(or_state* o, int n) {
    if (n < 2 || n > o->ng) return -1;
    mg_nlevel(o, n);
    mg_zero_pk_boundary(o, n);
    return 0;
}

/* rs_sweep (src/rs_sweep.F90:2-182): constant-coefficient implicit
 * smoothing of dq on interior nodes, with SGTSV and NRHS = 4 */
static void rs_sweep(or_state* o, int n, double* dq /* level n, stride 4 */) {
    LEVEL(n);
    int i, j, k, m, s, dir;
    int maxlen = ni > nj ? ni : nj;
    double *ax, *bx, *cx, *dx;
    if (nk > maxlen) maxlen = nk;
    ax = zalloc(maxlen); bx = zalloc(maxlen); cx = zalloc(maxlen); dx = zalloc(4 * maxlen);
    for (dir = 1; dir <= 3; dir++) {
        const double epn = o->p.ep[1][dir - 1];
        int len = (dir == 1 ? ni : dir == 2 ? nj : nk) - 2;
        int a, b, na = dir == 1 ? nj : ni, nb = dir == 3 ? nj : nk;
        for (b = 2; b <= nb - 1; b++)
            for (a = 2; a <= na - 1; a++) {
                /* line nodes: dir 1 (s+2, a, b), dir 2 (a, s+2, b), dir 3 (a, b, s+2) */
                for (s = 0; s < len; s++) {
                    int ii = dir == 1 ? s + 2 : a, jj = dir == 1 ? a : dir == 2 ? s + 2 : b, kk = dir == 3 ? s + 2 : b;
                    for (m = 0; m < 4; m++) dx[m * len + s] = dq[4 * IX(ii, jj, kk) + m];
                    ax[s] = -epn;
                    bx[s] = 1.0 + 2.0 * epn;
                    cx[s] = -epn;
                }
                or_gtsv(len, 4, ax + 1, bx, cx, dx, len);
                for (s = 0; s < len; s++) {
                    int ii = dir == 1 ? s + 2 : a, jj = dir == 1 ? a : dir == 2 ? s + 2 : b, kk = dir == 3 ? s + 2 : b;
                    for (m = 0; m < 4; m++) dq[4 * IX(ii, jj, kk) + m] = dx[m * len + s];
                }
            }
    }
    (void)i; (void)j; (void)k;
    free(ax); free(bx); free(cx); free(dx);
}

/* mg_prolong (src/mg_prolong.F90:1-281) */
static void mg_prolong(or_state* o, int n) {
    LEVEL(n);
    long N = (long)ni * nj * nk;
    double* dq = zalloc(4 * N);
    int i, j, k, m;
    const long ic = 1, ie = ni, iz = (long)ni * nj;
    const int icrs = o->p.icrs, jcrs = o->p.jcrs, kcrs = o->p.kcrs;
    /* coarse correction placed on the coincident fine nodes (:57-75) */
    for (k = 1; k <= o->kmg[n + 1]; k++)
        for (j = 1; j <= o->jmg[n + 1]; j++)
            for (i = 1; i <= o->img[n + 1]; i++) {
                int kk = (1 << kcrs) * k - kcrs, jj = (1 << jcrs) * j - jcrs, ii = (1 << icrs) * i - icrs;
                long ll = IX(ii, jj, kk), l = gidx(o, n + 1, i, j, k);
                for (m = 0; m < 4; m++) dq[4 * ll + m] = o->q[5 * l + m] - o->qold_mg[5 * l + m];
            }
    /* linear fill in k, then j, then i (:93-170) */
    if (kcrs == 1)
        for (k = 2; k <= nk - 1; k += 2)
            for (j = 1; j <= nj; j += 1 << jcrs)
                for (i = 1; i <= ni; i += 1 << icrs) {
                    long l = IX(i, j, k);
                    for (m = 0; m < 4; m++) dq[4 * l + m] = pt5 * dq[4 * (l + iz) + m] + pt5 * dq[4 * (l - iz) + m];
                }
    if (jcrs == 1)
        for (k = 1; k <= nk; k++)
            for (j = 2; j <= nj - 1; j += 2)
                for (i = 1; i <= ni; i += 1 << icrs) {
                    long l = IX(i, j, k);
                    for (m = 0; m < 4; m++) dq[4 * l + m] = pt5 * dq[4 * (l + ie) + m] + pt5 * dq[4 * (l - ie) + m];
                }
    if (icrs == 1)
        for (k = 1; k <= nk; k++)
            for (j = 1; j <= nj; j++)
                for (i = 2; i <= ni - 1; i += 2) {
                    long l = IX(i, j, k);
                    for (m = 0; m < 4; m++) dq[4 * l + m] = pt5 * dq[4 * (l + ic) + m] + pt5 * dq[4 * (l - ic) + m];
                }
    rs_sweep(o, n, dq);
    /* dq = 0 on the boundary planes (:218-270) */
    for (k = 1; k <= nk; k++)
        for (j = 1; j <= nj; j++)
            for (m = 0; m < 4; m++) dq[4 * IX(1, j, k) + m] = 0.0, dq[4 * IX(ni, j, k) + m] = 0.0;
    for (k = 1; k <= nk; k++)
        for (i = 1; i <= ni; i++)
            for (m = 0; m < 4; m++) dq[4 * IX(i, 1, k) + m] = 0.0, dq[4 * IX(i, nj, k) + m] = 0.0;
    for (j = 1; j <= nj; j++)
        for (i = 1; i <= ni; i++)
            for (m = 0; m < 4; m++) dq[4 * IX(i, j, 1) + m] = 0.0, dq[4 * IX(i, j, nk) + m] = 0.0;
    for (k = 1; k <= nk; k++)
        for (j = 1; j <= nj; j++)
            for (i = 1; i <= ni; i++) {
                long l = off + IX(i, j, k);
                for (m = 0; m < 4; m++) o->q[5 * l + m] = o->q[5 * l + m] + dq[4 * IX(i, j, k) + m];
            }
    free(dq);
}

int or_mg_prolong(or_state* o, int n) {
    if (n < 1 || n >= o->ng) return -1;
    mg_prolong(o, n);
    return 0;
}

/* mg_residual (src/mg_residual.F90:2-104): log10 of the L1 mean of the
 * pseudo-time update over fine interior nodes, loop order k, j, i */
/* the local sums resd_tmp of mg_residual (src/mg_residual.F90:57-67) */
static void mg_residual_sums(const or_state* o, double* resd) {
    LEVEL(1);
    int i, j, k, m;
    for (m = 0; m < 5; m++) resd[m] = 0.0;
    for (k = 2; k <= nk - 1; k++)
        for (j = 2; j <= nj - 1; j++)
            for (i = 2; i <= ni - 1; i++) {
                long l = off + IX(i, j, k);
                for (m = 0; m < 5; m++) resd[m] = resd[m] + fabs(o->q[5 * l + m] - o->qold_mg[5 * l + m]);
            }
}

int or_mg_residual(const or_state* o, double* residual) {
    LEVEL(1);
    double resd[5];
    double tnu = (double)((long)(ni - 2) * (nj - 2) * (nk - 2));
    int m;
    mg_residual_sums(o, resd);
    for (m = 0; m < 5; m++) residual[m] = log10(resd[m] / tnu);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Overset boundary interpolation (src/mg_bintp_mom.F90:2-172,          */
/* src/mg_bintp_tur.F90:2-182) for one zone on one rank: every donor     */
/* value is formed from the pre-update q (the send buffer of the         */
/* alltoallv) before any fringe node is written.                        */
/* ------------------------------------------------------------------ */
int or_set_overset(or_state* o, long n, const int* hosts, const double* coef, const int* nodes, int apply_mom) {
    free(o->ohosts); free(o->ocoef); free(o->onodes);
    o->ohosts = NULL; o->ocoef = NULL; o->onodes = NULL;
    o->nover = 0;
    o->apply_mom = apply_mom;
    if (n <= 0) return 0;
    LEVEL(1);
    for (long d = 0; d < n; d++) {
        const int* h = hosts + 3 * d;
        const int* t = nodes + 3 * d;
        if (h[0] < 1 || h[0] + 1 > ni || h[1] < 1 || h[1] + 1 > nj || h[2] < 1 || h[2] + 1 > nk) return -1;
        if (t[0] < 1 || t[0] > ni || t[1] < 1 || t[1] > nj || t[2] < 1 || t[2] > nk) return -1;
    }
    o->ohosts = (int*)malloc(sizeof(int) * 3 * n);
    o->onodes = (int*)malloc(sizeof(int) * 3 * n);
    o->ocoef = (double*)malloc(sizeof(double) * 3 * n);
    memcpy(o->ohosts, hosts, sizeof(int) * 3 * n);
    memcpy(o->onodes, nodes, sizeof(int) * 3 * n);
    memcpy(o->ocoef, coef, sizeof(double) * 3 * n);
    o->nover = n;
    return 0;
}

/* trilinear value of component m of q in the fine-level cell with lower
 * corner (ih, jh, kh) at weights (xco, yco, zco) (mg_bintp_mom.F90:85-131) */
static double trilin(const or_state* o, int ih, int jh, int kh, const double* w, int m) {
    LEVEL(1);
    const double xco = w[0], yco = w[1], zco = w[2];
    const long n1 = IX(ih, jh, kh), n2 = IX(ih + 1, jh, kh), n3 = IX(ih, jh + 1, kh);
    const long n4 = IX(ih, jh, kh + 1), n5 = IX(ih + 1, jh, kh + 1), n6 = IX(ih, jh + 1, kh + 1);
    const long n7 = IX(ih + 1, jh + 1, kh), n8 = IX(ih + 1, jh + 1, kh + 1);
    const double one = 1.0;
    const double c000 = (one - xco) * (one - yco) * (one - zco);
    const double c100 = xco * (one - yco) * (one - zco);
    const double c010 = (one - xco) * yco * (one - zco);
    const double c001 = (one - xco) * (one - yco) * zco;
    const double c101 = xco * (one - yco) * zco;
    const double c011 = (one - xco) * yco * zco;
    const double c110 = xco * yco * (one - zco);
    const double c111 = xco * yco * zco;
    const double* q = o->q;
    return q[5 * n1 + m] * c000 + q[5 * n2 + m] * c100 + q[5 * n3 + m] * c010 + q[5 * n4 + m] * c001 +
           q[5 * n5 + m] * c101 + q[5 * n6 + m] * c011 + q[5 * n7 + m] * c110 + q[5 * n8 + m] * c111;
}

static double bintp_value(const or_state* o, long d, int m) {
    return trilin(o, o->ohosts[3 * d], o->ohosts[3 * d + 1], o->ohosts[3 * d + 2], o->ocoef + 3 * d, m);
}

/* nu~ written into a fringe node, then its floor and xnut (mg_bintp_tur.F90:160-176) */
static void tur_fringe(or_state* o, long l, double v) {
    const double rei = 1.0 / o->p.ren, cv1c = 357.911;
    o->q[5 * l + 4] = v;
    if (o->p.des) {
        o->q[5 * l + 4] = dmax2(o->cvin, o->q[5 * l + 4]);
        const double chi = o->q[5 * l + 4] / rei;
        const double chic = chi * chi * chi;
        const double fv1 = chic / (chic + cv1c);
        o->xnut[l] = o->q[5 * l + 4] * fv1;
    }
}

/* mg_bintp_mom: q(1:4) at the fringe nodes (mg_bintp_mom.F90:59-165) */
static void bintp_mom(or_state* o) {
    if (o->nover == 0) return;
    LEVEL(1);
    double* buf = (double*)malloc(sizeof(double) * 4 * o->nover);
    for (long d = 0; d < o->nover; d++)
        for (int m = 0; m < 4; m++) buf[4 * d + m] = bintp_value(o, d, m);
    for (long d = 0; d < o->nover; d++) {
        const int* t = o->onodes + 3 * d;
        const long l = IX(t[0], t[1], t[2]);
        for (int m = 0; m < 4; m++) o->q[5 * l + m] = buf[4 * d + m];
    }
    free(buf);
}

/* mg_bintp_tur: nu~ at the fringe nodes, then the nu~ floor and xnut of
 * the SA model (mg_bintp_tur.F90:36-179, cv1^3 = 357.911 at :23) */
static void bintp_tur(or_state* o) {
    if (o->nover == 0) return;
    LEVEL(1);
    double* buf = (double*)malloc(sizeof(double) * o->nover);
    for (long d = 0; d < o->nover; d++) buf[d] = bintp_value(o, d, 4);
    for (long d = 0; d < o->nover; d++) {
        const int* t = o->onodes + 3 * d;
        tur_fringe(o, IX(t[0], t[1], t[2]), buf[d]);
    }
    free(buf);
}

int or_bintp(or_state* o, int which) {
    if (which == 0) bintp_mom(o);
    else bintp_tur(o);
    return 0;
}

/* ------------------------------------------------------------------ */
/* conservation (src/conservation.F90:1-174) on one rank and one zone   */
/* (igp = jgp = kgp = 0, init.F90:671): the contravariant flux U/J of   */
/* rhs_contra_j (src/rhs_contra_j.F90:12-27) is summed over the whole   */
/* i = 1 and i = img planes (j fastest, then k, as the send buffers are */
/* packed), and U/J on the outlet plane is scaled by fluxin / fluxout   */
/* before u, v, w are rebuilt there from the grid derivatives.          */
/* ------------------------------------------------------------------ */
static double contra(const or_state* o, const double* met, long l) {
    const double tmp = 1.0 / o->aj[l];
    return tmp * (met[3 * l] * o->q[5 * l + 1] + met[3 * l + 1] * o->q[5 * l + 2] + met[3 * l + 2] * o->q[5 * l + 3]);
}

static void conservation(or_state* o) {
    LEVEL(1);
    double fluxin = 0.0, fluxout = 0.0;
    for (int k = 1; k <= nk; k++)
        for (int j = 1; j <= nj; j++) fluxin = fluxin + contra(o, o->csi, IX(1, j, k));
    for (int k = 1; k <= nk; k++)
        for (int j = 1; j <= nj; j++) fluxout = fluxout + contra(o, o->csi, IX(ni, j, k));
    const double delta_e = fluxin - fluxout;
    for (int k = 1; k <= nk; k++)
        for (int j = 1; j <= nj; j++) {
            const long l = IX(ni, j, k);
            double u1 = contra(o, o->csi, l);
            const double u2 = contra(o, o->eta, l), u3 = contra(o, o->zet, l);
            u1 = u1 + delta_e * u1 / fluxout;
            const double a = o->aj[l];
            o->q[5 * l + 1] = a * (u1 * o->xc[l] + u2 * o->xe[l] + u3 * o->xz[l]);
            o->q[5 * l + 2] = a * (u1 * o->yc[l] + u2 * o->ye[l] + u3 * o->yz[l]);
            o->q[5 * l + 3] = a * (u1 * o->zc[l] + u2 * o->ze[l] + u3 * o->zz[l]);
        }
}

/* ------------------------------------------------------------------ */
/* Wall functions on the j walls: wf_eta (src/wf_mpi.F90:246-413) with  */
/* its constants (:26-30; ks = 0 switches the roughness shift off).     */
/* For every interior (i, k) of the wall j = 1 (jb = 1) or j = jmg      */
/* (jb = 2) the friction velocity u_tau solves                          */
/*   f(u) = (ln(Re y u) / kappa + B) u - u0   if Re y u > 11.3,          */
/*   f(u) = Re y u^2 - u0                      otherwise,                 */
/*   f = 0 where |u Re y - 11.3| < 1e-4 (funcv, :604-631),               */
/* with u0 the wall-parallel speed and y the wall distance two nodes    */
/* off the wall; the first node off the wall then gets the log-law (or  */
/* linear) speed in the direction of the third node's velocity and      */
/* nu~ = kappa y+ / Re.                                                 */
/* ------------------------------------------------------------------ */
static const double wf_kapa = 0.41, wf_blog = 5.45, wf_yplm = 11.30, wf_qmin = 1.0e-08;

static double wf_f(double x, double u0, double yp0, double re0) {
    double f;
    if (re0 * yp0 * x > 11.3) f = (1.0 * log(re0 * yp0 * x) / wf_kapa + wf_blog - 0.0) * x - u0;
    else f = re0 * (x * x) * yp0 - u0;
    if (fabs(x * re0 * yp0 - 11.3) < 1.0e-4) f = 0.0;
    return f;
}

static double wf_df(double x, double yp0, double re0) {
    if (re0 * yp0 * x > 11.3) return log(re0 * yp0 * x) / wf_kapa + wf_blog + 1.0 / wf_kapa;
    return re0 * 2.0 * x * yp0;
}

/* The reference's solve: newt (:634-720) for one unknown, i.e. ludcmp /
 * lubksb reduce to p = -f / f', with the backtracking line search lnsrch
 * (:765-850); the globally convergent Newton method of Press et al.,
 * Numerical Recipes (2nd ed., section 9.7).  x is the start and the result.
 * Returns -1 where the reference stops the program (a zero derivative in
 * ludcmp, a non-descent direction in lnsrch). */
static int wf_newt(double* xp, double u0, double yp0, double re0) {
    const double TOLF = 1.e-4, TOLMIN = 1.e-6, TOLX = 1.e-7, STPMX = 100., ALF = 1.e-3;
    double x = *xp;
    double fvec = wf_f(x, u0, yp0, re0);
    double f = 0.5 * (0.0 + fvec * fvec);
    double test = fabs(fvec);
    if (test < .01 * TOLF) return 0;
    double sum = 0.0 + x * x;
    const double stpmax = STPMX * dmax2(sqrt(sum), 1.0);
    for (int its = 1; its <= 200; its++) {
        fvec = wf_f(x, u0, yp0, re0);
        const double fjac = wf_df(x, yp0, re0);
        const double g = 0.0 + fjac * fvec;
        const double xold = x, fold = f;
        double p = -fvec;
        if (fabs(fjac) == 0.0) return -1;             /* 'singular matrix in ludcmp' */
        p = p / fjac;
        /* lnsrch */
        int check = 0;
        sum = sqrt(0.0 + p * p);
        if (sum > stpmax) p = p * stpmax / sum;
        const double slope = 0.0 + g * p;
        if (slope >= 0.) return -1;                   /* 'roundoff problem in lnsrch' */
        const double alamin = TOLX / (fabs(p) / dmax2(fabs(xold), 1.));
        double alam = 1., alam2 = 2., f2 = 2., tmplam = 0.0;
        for (;;) {
            x = xold + alam * p;
            fvec = wf_f(x, u0, yp0, re0);
            f = 0.5 * (0.0 + fvec * fvec);
            if (alam < alamin) {
                x = xold;
                check = 1;
                break;
            } else if (f <= fold + ALF * alam * slope) {
                break;
            }
            if (alam == 1.) {
                tmplam = -slope / (2. * (f - fold - slope));
            } else {
                const double rhs1 = f - fold - alam * slope;
                const double rhs2 = f2 - fold - alam2 * slope;
                const double a = (rhs1 / (alam * alam) - rhs2 / (alam2 * alam2)) / (alam - alam2);
                const double b = (-(alam2 * rhs1 / (alam * alam)) + alam * rhs2 / (alam2 * alam2)) / (alam - alam2);
                if (a == 0.) {
                    tmplam = -slope / (2. * b);
                } else {
                    const double disc = b * b - 3. * a * slope;
                    if (disc < 0.) tmplam = .5 * alam;
                    else if (b <= 0.) tmplam = (-b + sqrt(disc)) / (3. * a);
                    else tmplam = -slope / (b + sqrt(disc));
                }
                if (tmplam > .5 * alam) tmplam = .5 * alam;
            }
            alam2 = alam;
            f2 = f;
            alam = dmax2(tmplam, .1 * alam);
        }
        /* back in newt */
        if (fabs(fvec) < TOLF) break;
        if (check) break;
        if (fabs(x - xold) / dmax2(fabs(x), 1.) < TOLX) break;
    }
    *xp = x;
    return 0;
}

/* The same root converged to rounding: f is increasing in u > 0 (both
 * branches), convex on the log branch and jumps up at Re y u = 11.3, so
 * the root is sqrt(u0 / (Re y)) when that lies on the quadratic branch,
 * the jump point when f changes sign across it, and otherwise the limit
 * of Newton's iteration started right of the root on the log branch
 * (monotone for a convex increasing function).  Every node is solved on
 * its own; the result does not depend on a start. */
static double wf_root(double u0, double yp0, double re0) {
    if (!(u0 > 0.0)) return u0 == 0.0 ? 0.0 : u0 - u0;
    const double c = re0 * yp0;
    const double us = sqrt(u0 / c);
    if (c * us <= 11.3) return us;
    const double lo = 11.3 / c;
    if (log(c * lo) / wf_kapa * lo + wf_blog * lo - u0 >= 0.0) return lo;
    double x = 2.0 * lo;
    for (int t = 0; t < 200 && (log(c * x) / wf_kapa + wf_blog) * x - u0 < 0.0; t++) x = 2.0 * x;
    for (int t = 0; t < 100; t++) {
        const double lg = log(c * x);
        const double fx = (lg / wf_kapa + wf_blog) * x - u0;
        const double step = fx / (lg / wf_kapa + wf_blog + 1.0 / wf_kapa);
        const double xn = x - step;
        if (!(xn < x) || xn <= lo) break;
        x = xn;
    }
    return x;
}

double or_wf_root(double u0, double yp0, double re0) { return wf_root(u0, yp0, re0); }

int or_wf_newt(double* x, double u0, double yp0, double re0) { return wf_newt(x, u0, yp0, re0); }

static int wf_eta(or_state* o, int jb) {
    LEVEL(1);
    const int j = jb == 1 ? 1 : nj, j1 = jb == 1 ? 2 : nj - 1, j2 = jb == 1 ? 3 : nj - 2;
    const double re0 = o->p.ren;
    double* ut0 = (double*)malloc(sizeof(double) * (size_t)ni * nk);
    double ut = 0.03;  /* the reference carries the start from node to node */
    const double* q = o->q;
    for (int k = 2; k <= nk - 1; k++)
        for (int i = 2; i <= ni - 1; i++) {
            const long l = IX(i, j, k), l2 = IX(i, j2, k);
            const double* e = o->eta + 3 * l;
            const double vmod = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
            const double qnor = (q[5 * l2 + 1] * e[0] + q[5 * l2 + 2] * e[1] + q[5 * l2 + 3] * e[2]) / vmod;
            double qpar = q[5 * l2 + 1] * q[5 * l2 + 1] + q[5 * l2 + 2] * q[5 * l2 + 2] + q[5 * l2 + 3] * q[5 * l2 + 3];
            qpar = sqrt(qpar - qnor * qnor);
            const double u0 = qpar, yp0 = o->wd[l2];
            if (o->wf_newton) {
                if (wf_newt(&ut, u0, yp0, re0)) {
                    free(ut0);
                    return -1;
                }
            } else {
                ut = wf_root(u0, yp0, re0);
            }
            ut0[(i - 1) + (long)ni * (k - 1)] = ut;
        }
    const double* x = o->x;
    const double* y = o->y;
    const double* z = o->z;
    for (int k = 2; k <= nk - 1; k++)
        for (int i = 2; i <= ni - 1; i++) {
            const long l1 = IX(i, j1, k), l2 = IX(i, j2, k);
            const double* q2v = q + 5 * l2;
            double* q1v = o->q + 5 * l1;
            const double qpar = q2v[1] * q2v[1] + q2v[2] * q2v[2] + q2v[3] * q2v[3];
            const double u = ut0[(i - 1) + (long)ni * (k - 1)];
            if (qpar > wf_qmin) {
                const double yplus = o->wd[l1] * o->p.ren * u;
                const double deltb = 0.0;
                double qw;
                if (yplus > wf_yplm) qw = (log(yplus) / wf_kapa + wf_blog - deltb) * u;
                else qw = yplus * u;
                long lu = IX(i + 1, j, k), ld = IX(i - 1, j, k);
                const double xc = pt5 * (x[lu] - x[ld]), yc = pt5 * (y[lu] - y[ld]), zc = pt5 * (z[lu] - z[ld]);
                lu = IX(i, j, k + 1);
                ld = IX(i, j, k - 1);
                const double xz = pt5 * (x[lu] - x[ld]), yz = pt5 * (y[lu] - y[ld]), zz = pt5 * (z[lu] - z[ld]);
                const double sg11 = sqrt(xc * xc + yc * yc + zc * zc);
                const double sg33 = sqrt(xz * xz + yz * yz + zz * zz);
                const double* cs = o->csi + 3 * l2;
                const double* zt = o->zet + 3 * l2;
                double qc1 = (cs[0] * q2v[1] + cs[1] * q2v[2] + cs[2] * q2v[3]) * sg11;
                double qc2 = (zt[0] * q2v[1] + zt[1] * q2v[2] + zt[2] * q2v[3]) * sg33;
                const double cgama = (xc * xz + yc * yz + zc * zz) / (sg11 * sg33);
                double fbeta = sqrt(qc1 * qc1 + qc2 * qc2 + 2. * qc1 * qc2 * cgama);
                fbeta = ((qc1 + qc2 * cgama) / fbeta);
                fbeta = dmin2(fbeta, 1.0);
                fbeta = dmax2(fbeta, -1.0);
                if (qc2 > 0.0) fbeta = acos(fbeta);
                else fbeta = -acos(fbeta);
                const double qstar = qw;
                qc2 = qstar * sin(fbeta) / sin(acos(cgama));
                qc1 = qstar * cos(fbeta) - qc2 * cgama;
                q1v[1] = qc1 * xc / sg11 + qc2 * xz / sg33;
                q1v[2] = qc1 * yc / sg11 + qc2 * yz / sg33;
                q1v[3] = qc1 * zc / sg11 + qc2 * zz / sg33;
                q1v[4] = wf_kapa * yplus / re0;
            } else {
                q1v[1] = 0.0;
                q1v[2] = 0.0;
                q1v[3] = 0.0;
                q1v[4] = 0.1 / re0;
            }
        }
    free(ut0);
    return 0;
}

int or_set_hooks(or_state* o, int hooks, int wf_newton) {
    if (hooks & ~(OR_HOOK_CONSERVATION | OR_HOOK_WALL_FUNCTIONS)) return -1;
    o->hooks = hooks;
    o->wf_newton = wf_newton;
    return 0;
}

/* mg_driver (src/mg_driver.F90:2-323) with mg_ck_conver (:335-405) */
int or_mg_driver(or_state* o, int* itc_out, double* resid_hist) {
    const int ng = o->ng, ns = 1;
    long n1 = o->nl[1], l;
    int m, it, n, itr, decide_recalc_rh = 0;
    if (o->img[ng] < 5 || o->jmg[ng] < 5 || o->kmg[ng] < 5) return -1;
    o->itc = 0;
    /* qnm1 <- qn, qn <- q on the fine level (:48-51) */
    for (l = 0; l < 5 * n1; l++) {
        o->qnm1[l] = o->qn[l];
        o->qn[l] = o->q[l];
    }
    for (m = 0; m < 5; m++) {
        mg_inject(o, o->qn, 5, m);
        mg_inject(o, o->qnm1, 5, m);
    }
    if (o->hooks & OR_HOOK_CONSERVATION) conservation(o); /* :64-75 */
    for (it = 1; it <= o->p.itm; it++) {
        double residual[5], eomax, ermax;
        int converged = 0;
        o->itc++;
        memcpy(o->qold_mg, o->q, 5 * n1 * sizeof(double));
        for (n = ns; n <= ng; n++) {
            if (n > 1) {
                mg_nlevel(o, n);
                mg_zero_pk_boundary(o, n);
            }
            for (itr = 1; itr <= o->p.iter[n - 1]; itr++) {
                if (n != ng) decide_recalc_rh = (itr != o->p.iter[n - 1]) ? 0 : 1;
                if (o->p.des && n == ns) {
                    des_eddy(o);
                    bintp_tur(o);
                }
                solver_daf(o, n, decide_recalc_rh, itr);
                /* wall functions of zone 1 on the j walls (:218-224) */
                if (n == ns && (o->hooks & OR_HOOK_WALL_FUNCTIONS))
                    if (wf_eta(o, 1) || wf_eta(o, 2)) return -2;
                if (n == ns && o->apply_mom) bintp_mom(o);
            }
            mg_zero_pk_boundary(o, n);
        }
        for (n = ng - 1; n >= ns; n--) mg_prolong(o, n);
        bcond_fm(o, o->q);
        or_mg_residual(o, residual);
        if (resid_hist)
            for (m = 0; m < 5; m++) resid_hist[5 * (it - 1) + m] = residual[m];
        if (o->itc == 1)
            for (m = 0; m < 5; m++) o->init_resd[m] = residual[m];
        eomax = dmax2(dmax2(residual[1], residual[2]), residual[3]);
        ermax = dmax2(dmax2(residual[1] - o->init_resd[1], residual[2] - o->init_resd[2]), residual[3] - o->init_resd[3]);
        if (o->itc > o->p.it_min && ermax < o->p.er_min) converged = 1;
        if (o->itc > o->p.it_min && eomax < o->p.eo_min) converged = 1;
        if (converged) break;
    }
    if (itc_out) *itc_out = o->itc;
    return 0;
}

/* ------------------------------------------------------------------ */
/* Several zones, one rank each in the reference (nzone > 1), coupled by */
/* a cross-zone donor table.  Each rank runs mg_driver; the ranks meet  */
/* only in the overset exchanges (mg_bintp_tur after des_eddy,          */
/* mg_bintp_mom after the fine pass, mg_driver.F90:182,241-244) and in  */
/* the norm reduction of mg_residual (mpi_reduce of the local sums to   */
/* the root, divided by the root zone's interior count,                  */
/* mg_residual.F90:36,71-82).  Running every zone's work between two    */
/* exchanges in zone order is therefore the same computation.           */
/* hosts, nodes: 4 ints per donor (i, j, k, zone), 1-based, as          */
/* hosts(1:4,.)/nodes(1:4,.) of global_osg.F90:18-24.                    */
/* ------------------------------------------------------------------ */
static int group_check(or_state* const* zs, int nz, long nd, const int* hosts, const int* nodes) {
    for (long d = 0; d < nd; d++) {
        const int* h = hosts + 4 * d;
        const int* t = nodes + 4 * d;
        if (h[3] < 1 || h[3] > nz || t[3] < 1 || t[3] > nz) return -1;
        const or_state* a = zs[h[3] - 1];
        const or_state* b = zs[t[3] - 1];
        if (h[0] < 1 || h[0] + 1 > a->img[1] || h[1] < 1 || h[1] + 1 > a->jmg[1] || h[2] < 1 || h[2] + 1 > a->kmg[1])
            return -1;
        if (t[0] < 1 || t[0] > b->img[1] || t[1] < 1 || t[1] > b->jmg[1] || t[2] < 1 || t[2] > b->kmg[1]) return -1;
    }
    return 0;
}

static void group_bintp(or_state* const* zs, long nd, const int* hosts, const double* coef, const int* nodes,
                        int tur) {
    const int nc = tur ? 1 : 4, m0 = tur ? 4 : 0;
    double* buf = (double*)malloc(sizeof(double) * (nd > 0 ? nd : 1) * nc);
    for (long d = 0; d < nd; d++) {
        const int* h = hosts + 4 * d;
        for (int c = 0; c < nc; c++) buf[nc * d + c] = trilin(zs[h[3] - 1], h[0], h[1], h[2], coef + 3 * d, m0 + c);
    }
    for (long d = 0; d < nd; d++) {
        const int* t = nodes + 4 * d;
        or_state* o = zs[t[3] - 1];
        const int ni = o->img[1], nj = o->jmg[1];
        const long l = IX(t[0], t[1], t[2]);
        if (tur) tur_fringe(o, l, buf[d]);
        else
            for (int m = 0; m < 4; m++) o->q[5 * l + m] = buf[4 * d + m];
    }
    free(buf);
}

int or_group_bintp(or_state* const* zs, int nz, long nd, const int* hosts, const double* coef, const int* nodes,
                   int which) {
    if (group_check(zs, nz, nd, hosts, nodes)) return -1;
    group_bintp(zs, nd, hosts, coef, nodes, which != 0);
    return 0;
}

int or_group_mg_driver(or_state* const* zs, int nz, long nd, const int* hosts, const double* coef, const int* nodes,
                       int apply_mom, int* itc_out, double* resid_hist) {
    if (nz < 1) return -1;
    const or_state* z0 = zs[0];
    const int ng = z0->ng, ns = 1;
    int z, m, it, n, itr, decide_recalc_rh = 0, itc = 0;
    double init_resd[5] = {0, 0, 0, 0, 0};
    for (z = 0; z < nz; z++) {
        const or_state* o = zs[z];
        if (o->ng != ng || o->img[ng] < 5 || o->jmg[ng] < 5 || o->kmg[ng] < 5) return -1;
        for (n = 0; n < ng; n++)
            if (o->p.iter[n] != z0->p.iter[n]) return -1;
        if (o->p.des != z0->p.des || o->p.itm != z0->p.itm) return -1;
    }
    if (group_check(zs, nz, nd, hosts, nodes)) return -1;
    for (z = 0; z < nz; z++) {
        or_state* o = zs[z];
        const long n1 = o->nl[1];
        o->itc = 0;
        for (long l = 0; l < 5 * n1; l++) {
            o->qnm1[l] = o->qn[l];
            o->qn[l] = o->q[l];
        }
        for (m = 0; m < 5; m++) {
            mg_inject(o, o->qn, 5, m);
            mg_inject(o, o->qnm1, 5, m);
        }
    }
    const double tnu = (double)((long)(z0->img[1] - 2) * (z0->jmg[1] - 2) * (z0->kmg[1] - 2));
    for (it = 1; it <= z0->p.itm; it++) {
        double resd[5] = {0, 0, 0, 0, 0}, residual[5], eomax, ermax;
        int converged = 0;
        itc++;
        for (z = 0; z < nz; z++) {
            zs[z]->itc = itc;
            memcpy(zs[z]->qold_mg, zs[z]->q, 5 * zs[z]->nl[1] * sizeof(double));
        }
        for (n = ns; n <= ng; n++) {
            if (n > 1)
                for (z = 0; z < nz; z++) {
                    mg_nlevel(zs[z], n);
                    mg_zero_pk_boundary(zs[z], n);
                }
            for (itr = 1; itr <= z0->p.iter[n - 1]; itr++) {
                if (n != ng) decide_recalc_rh = (itr != z0->p.iter[n - 1]) ? 0 : 1;
                if (z0->p.des && n == ns) {
                    for (z = 0; z < nz; z++) des_eddy(zs[z]);
                    group_bintp(zs, nd, hosts, coef, nodes, 1);
                }
                for (z = 0; z < nz; z++) solver_daf(zs[z], n, decide_recalc_rh, itr);
                if (n == ns && apply_mom) group_bintp(zs, nd, hosts, coef, nodes, 0);
            }
            for (z = 0; z < nz; z++) mg_zero_pk_boundary(zs[z], n);
        }
        for (z = 0; z < nz; z++) {
            for (n = ng - 1; n >= ns; n--) mg_prolong(zs[z], n);
            bcond_fm(zs[z], zs[z]->q);
        }
        for (z = 0; z < nz; z++) {
            double part[5];
            mg_residual_sums(zs[z], part);
            for (m = 0; m < 5; m++) resd[m] = resd[m] + part[m];
        }
        for (m = 0; m < 5; m++) residual[m] = log10(resd[m] / tnu);
        if (resid_hist)
            for (m = 0; m < 5; m++) resid_hist[5 * (it - 1) + m] = residual[m];
        if (itc == 1)
            for (m = 0; m < 5; m++) init_resd[m] = residual[m];
        eomax = dmax2(dmax2(residual[1], residual[2]), residual[3]);
        ermax = dmax2(dmax2(residual[1] - init_resd[1], residual[2] - init_resd[2]), residual[3] - init_resd[3]);
        if (itc > z0->p.it_min && ermax < z0->p.er_min) converged = 1;
        if (itc > z0->p.it_min && eomax < z0->p.eo_min) converged = 1;
        if (converged) break;
    }
    if (itc_out) *itc_out = itc;
    return 0;
}

/* ------------------------------------------------------------------ */
/* actuator-line tables (deep copies)                                   */
/* ------------------------------------------------------------------ */
static void* dup(const void* p, size_t bytes) {
    void* q = malloc(bytes ? bytes : 1);
    if (bytes) memcpy(q, p, bytes);
    return q;
}

int or_set_turbines(or_state* o, int n_tur, const desalm_turbine* tur, int n_polar, const desalm_polar* pol,
                    double length_scale, double vel_infty) {
    for (int t = 0; t < o->n_tur; t++) {
        desalm_turbine* T = &o->tur[t];
        free((void*)T->node_indexes); free((void*)T->node_indexes_pos); free((void*)T->disk_dat);
        free((void*)T->polar_before); free((void*)T->polar_after); free((void*)T->disk_nodes);
        free((void*)T->ker_nodes); free((void*)T->ker_pos); free((void*)T->ker_dat);
    }
    for (int p = 0; p < o->n_polar; p++) {
        free((void*)o->pol[p].angle); free((void*)o->pol[p].cl); free((void*)o->pol[p].cd);
    }
    free(o->tur); free(o->pol);
    o->tur = NULL; o->pol = NULL; o->n_tur = 0; o->n_polar = 0;
    if (n_tur <= 0) return 0;
    const long jk = (long)o->jmg[1] * o->kmg[1];
    o->tur = (desalm_turbine*)calloc(n_tur, sizeof(desalm_turbine));
    o->pol = (desalm_polar*)calloc(n_polar > 0 ? n_polar : 1, sizeof(desalm_polar));
    for (int t = 0; t < n_tur; t++) {
        desalm_turbine* T = &o->tur[t];
        *T = tur[t];
        const long na = (long)T->nelems * T->nangles;
        T->node_indexes = (const int32_t*)dup(tur[t].node_indexes, sizeof(int32_t) * 3 * na);
        T->node_indexes_pos = (const double*)dup(tur[t].node_indexes_pos, sizeof(double) * 27 * na);
        T->disk_dat = (const double*)dup(tur[t].disk_dat, sizeof(double) * 8 * jk);
        T->polar_before = (const int32_t*)dup(tur[t].polar_before, sizeof(int32_t) * jk);
        T->polar_after = (const int32_t*)dup(tur[t].polar_after, sizeof(int32_t) * jk);
        T->disk_nodes = (const int32_t*)dup(tur[t].disk_nodes, sizeof(int32_t) * 3 * T->n_disk);
        T->ker_nodes = (const int32_t*)dup(tur[t].ker_nodes, sizeof(int32_t) * 3 * T->n_ker);
        T->ker_pos = (const double*)dup(tur[t].ker_pos, sizeof(double) * 3 * T->n_ker);
        T->ker_dat = (const double*)dup(tur[t].ker_dat, sizeof(double) * 3 * T->n_ker);
    }
    for (int p = 0; p < n_polar; p++) {
        o->pol[p].n = pol[p].n;
        o->pol[p].angle = (const double*)dup(pol[p].angle, sizeof(double) * pol[p].n);
        o->pol[p].cl = (const double*)dup(pol[p].cl, sizeof(double) * pol[p].n);
        o->pol[p].cd = (const double*)dup(pol[p].cd, sizeof(double) * pol[p].n);
    }
    o->n_tur = n_tur;
    o->n_polar = n_polar;
    o->length_scale = length_scale;
    o->vel_infty = vel_infty;
    return 0;
}
