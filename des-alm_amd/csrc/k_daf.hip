// k_daf.hip -- diagonalized approximate-factorization implicit operator.
//
// rhs_modal_matrices + rhs_diag_solver (src/rhs_modal_matrices.F90,
// src/rhs_diag_solver.F90).  Each direction D is one kernel, k_daf_sweep<D>:
// it loads the rows of its lines, applies the modal transform that precedes
// the sweep (S and Ma^-1 for i, N1^-1 = Mb^-1 Ma for j, N2^-1 = Mc^-1 Mb for k)
// as the rows arrive, solves the four scalar tridiagonal systems of every
// line and writes the solution; the k sweep applies Mc and adds dq to q
// instead (rhs_diag_solver.F90:624-633, solver_daf.F90:409-418).  The
// eigenvector matrices are rebuilt in registers from the metrics and dtau of
// each node instead of being stored (the reference keeps 67 words per node),
// and the line solves keep every row on chip.
//
// Line solve.  Thomas elimination is a pair of first-order recurrences,
//   c'_s = c_s / (b_s - a_s c'_(s-1)),  d'_s = (d_s - a_s d'_(s-1)) / den_s,
// and the back substitution x_s = d'_s - c'_s x_(s+1).  A workgroup owns TI
// lines and cuts each into P segments of at most R consecutive rows; the
// thread of a segment holds its rows in registers.  Each recurrence, seen as
// a map from the value entering a segment to the value leaving it, is a
// Moebius map (c') or an affine map (d', x) that the thread composes over its
// rows; one thread per (line, system) then walks the P segment maps in LDS to
// find the value entering every segment, and every thread runs the exact
// sequential recurrence over its own rows from there.  So each row is
// eliminated with the recurrence LAPACK SGTSV applies when it does not
// interchange rows, and the values at segment starts differ from a one-thread
// sweep only by rounding.  The three systems with distinct matrices are
// eliminated once each: systems 1 and 2 (eigenvalue 0) share theirs.
//
// SGTSV (src/sgtsv.f:89-259) interchanges rows where |den_s| < |a_(s+1)|;
// every sweep counts those rows in F.pivots (systems 1 and 2 each), so a
// caller sees when the pivot-free elimination left SGTSV's row order, and
// records the first row whose pivot den_s is zero or not finite (1-based, in
// line order) in F.zero_pivot, the INFO that SGTSV would return.
#include <algorithm>

#include "desalm_internal.h"

namespace desalm {

namespace {

// ---- modal matrices (rhs_modal_matrices.F90:112-264), applied in registers ----
// dtau-dependent factors of a node: sigma = sqrt(beta alpha) with
// alpha = 1 + e_source 1.5 dtau / delti, and theta = sqrt(beta / alpha) =
// beta / sigma (rhs_modal_matrices.F90:82-84)
struct NodeDt {
    double dt, alpha, sigma, isigma, theta;
};
__device__ __forceinline__ NodeDt node_dt(const Consts& C, double dt) {
    NodeDt n;
    n.dt = dt;
    n.alpha = 1.0 + C.e_source * 1.5 * dt / C.delti;
    n.sigma = sqrt(C.beta * n.alpha);
    n.isigma = 1.0 / n.sigma;
    n.theta = C.beta * n.isigma;
    return n;
}

// v <- M v for the Ma-type matrix of metric row (m1, m2, m3):
//   row 0 = (0, 0, sigma, sigma); rows 1-3 = (t1/s, t2/(s sg), m/sg, -m/sg)
//   with sg = |m|, s = sqrt(2 |m|^2 + (m2 - m3)^2)
__device__ __forceinline__ void mul_m(double m1, double m2, double m3, double sigma, double v[4]) {
    const double g = m1 * m1 + m2 * m2 + m3 * m3;
    const double isg = 1.0 / sqrt(g);
    const double is = 1.0 / sqrt(2.0 * g + (m2 - m3) * (m2 - m3));
    const double w0 = v[0] * is, w1 = v[1] * (is * isg), u = (v[2] - v[3]) * isg;
    const double dm = m3 - m2;
    const double o0 = sigma * (v[2] + v[3]);
    const double o1 = -(m2 + m3) * w0 + m1 * dm * w1 + m1 * u;
    const double o2 = m1 * w0 + (g + m2 * dm) * w1 + m2 * u;
    const double o3 = m1 * w0 - (g - m3 * dm) * w1 + m3 * u;
    v[0] = o0; v[1] = o1; v[2] = o2; v[3] = o3;
}

// v <- M^-1 v for the same matrix, given g = |m|^2 and sg = |m| (the inverse
// in closed form, d = |m|^2 + 2 m2 m3 + m1^2)
__device__ __forceinline__ void mul_minv(double m1, double m2, double m3, double g, double sg, double isigma,
                                         double v[4]) {
    const double m1s = m1 * m1, m2s = m2 * m2, m3s = m3 * m3;
    const double s = sqrt(2.0 * g + (m2 - m3) * (m2 - m3));
    const double id = 1.0 / (g + 2.0 * m2 * m3 + m1s);
    const double isg = 1.0 / sg;
    const double r1 = m2 + m3, r2 = m2 - m3, r3 = m2s - m3s;
    const double o0 = s * id * (m1 * (v[2] + v[3]) - r1 * v[1]);
    const double o1 = s * isg * id *
                      (m1 * (m3 - m2) * v[1] + (m2 * m3 + m1s + m3s) * v[2] - (m2 * m3 + m1s + m2s) * v[3]);
    const double t = 0.5 * isigma * v[0];
    const double p = 0.5 * isg * id *
                     (m1 * (2.0 * g + 2.0 * m2 * m3 - m2s - m3s) * v[1] + (g * r1 + m1s * r2 + m3 * r3) * v[2] +
                      (g * r1 - m1s * r2 - m2 * r3) * v[3]);
    v[0] = o0; v[1] = o1; v[2] = t + p; v[3] = t - p;
}

// S (-dtau J rh) at global node g (rhs_diag_solver.F90:56-72): beta on the
// pressure row, 1 / alpha on the velocity rows
__device__ __forceinline__ void scaled_rh_vals(const Consts& C, double aj, const double rh[4], const NodeDt& nd,
                                               double v[4]) {
    const double f = -nd.dt * aj;
    const double ia = 1.0 / nd.alpha;
    v[0] = C.beta * (f * rh[0]);
#pragma unroll
    for (int m = 1; m < 4; m++) v[m] = ia * (f * rh[m]);
}
__device__ __forceinline__ void scaled_rh(const Fields& F, const Consts& C, long g, const NodeDt& nd, double v[4]) {
    const double rh[4] = {F.rh[0][g], F.rh[1][g], F.rh[2][g], F.rh[3][g]};
    scaled_rh_vals(C, F.aj[g], rh, nd, v);
}

// the metric row of direction D at node g
template <int D>
__device__ __forceinline__ void metric_row(const Fields& F, long g, double& m1, double& m2, double& m3) {
    const double* const* mt = D == 0 ? F.csi : D == 1 ? F.eta : F.zet;
    m1 = mt[0][g]; m2 = mt[1][g]; m3 = mt[2][g];
}

// The modal transform that precedes sweep D at node g, applied to v
// (Ma^-1; N1^-1 = Mb^-1 Ma; N2^-1 = Mc^-1 Mb, applied factor by factor),
// and dtau times the spectral radius of direction D at g,
// dtau theta |grad xi_D| (rhs_diag_solver.F90:149-160).
// D = 0 builds its input S (-dtau J rh) here; D = 1, 2 take v as given.
template <int D>
__device__ __forceinline__ double transform_row(const Fields& F, const Consts& C, long g, double v[4]) {
    const NodeDt nd = node_dt(C, F.dtau[g]);
    if (D == 0) scaled_rh(F, C, g, nd, v);
    if (D > 0) {
        double a1, a2, a3;
        metric_row<D - 1>(F, g, a1, a2, a3);
        mul_m(a1, a2, a3, nd.sigma, v);
    }
    double m1, m2, m3;
    metric_row<D>(F, g, m1, m2, m3);
    const double gg = m1 * m1 + m2 * m2 + m3 * m3;
    const double sg = sqrt(gg);
    mul_minv(m1, m2, m3, gg, sg, nd.isigma, v);
    return nd.dt * (nd.theta * sg);
}

// the i sweep's transform (S then Ma^-1) and dtau spr from values already
// loaded: dtau, J, rh and the csi row of one node
__device__ __forceinline__ double transform_vals0(const Consts& C, double dt, double aj, const double rh[4], double m1,
                                                  double m2, double m3, double v[4]) {
    const NodeDt nd = node_dt(C, dt);
    scaled_rh_vals(C, aj, rh, nd, v);
    const double gg = m1 * m1 + m2 * m2 + m3 * m3;
    const double sg = sqrt(gg);
    mul_minv(m1, m2, m3, gg, sg, nd.isigma, v);
    return nd.dt * (nd.theta * sg);
}
__device__ __forceinline__ double sd_vals(const Consts& C, double dt, double m1, double m2, double m3) {
    const NodeDt nd = node_dt(C, dt);
    return nd.dt * (nd.theta * sqrt(m1 * m1 + m2 * m2 + m3 * m3));
}

// dtau times the spectral radius of direction D alone
template <int D>
__device__ __forceinline__ double row_sd(const Fields& F, const Consts& C, long g) {
    const NodeDt nd = node_dt(C, F.dtau[g]);
    double m1, m2, m3;
    metric_row<D>(F, g, m1, m2, m3);
    return nd.dt * (nd.theta * sqrt(m1 * m1 + m2 * m2 + m3 * m3));
}

// the transform of sweep D alone (boundary vectors)
template <int D>
__device__ __forceinline__ void pre_transform(const Fields& F, const Consts& C, long g, double v[4]) {
    const NodeDt nd = node_dt(C, F.dtau[g]);
    if (D > 0) {
        double a1, a2, a3;
        metric_row<D - 1>(F, g, a1, a2, a3);
        mul_m(a1, a2, a3, nd.sigma, v);
    }
    double m1, m2, m3;
    metric_row<D>(F, g, m1, m2, m3);
    const double gg = m1 * m1 + m2 * m2 + m3 * m3;
    mul_minv(m1, m2, m3, gg, sqrt(gg), nd.isigma, v);
}

// v <- Mc v at global node g (rhs_diag_solver.F90:624-633)
__device__ __forceinline__ void apply_mc(const Consts& C, double dtau, double z1, double z2, double z3, double v[4]) {
    const NodeDt nd = node_dt(C, dtau);
    mul_m(z1, z2, z3, nd.sigma, v);
}

template <typename S>
struct SweepIO {
    const S* xin[4];  // j and k sweeps: the previous sweep's solution (level-local)
    S* xout[4];       // i and j sweeps: this sweep's solution (level-local)
};

struct SweepShape {
    int n;        // rows per line
    int P;        // segments per line
    int TI;       // lines per workgroup
    int nI;       // i positions of the j and k lines
    long nlines;
};

// values of one segment slot in LDS (the largest phase: 3 Moebius maps, 12
// values) rounded up to an odd count, and a line's slots padded to an odd
// count: the threads of adjacent segments, and the scanning threads of
// adjacent lines, then fall on different LDS banks
constexpr int kSlot = 13;
__host__ __device__ constexpr int line_stride(int P) { return (P * kSlot) | 1; }

// tridiagonal row of matrix mi (eigenvalue 0, +1, -1) from sd = dtau spr of
// the row's node and its two neighbours (rhs_diag_solver.F90:149-189); the
// characteristic-outlet row (ext) couples the +1 eigenvalue upwind only
// (:207-221)
template <typename S>
__device__ __forceinline__ void row_coef(int mi, S sdm, S sd0, S sdp, S h, S h2, S hsq, S epd, bool outlet, S& a,
                                         S& b, S& c) {
    const S lam = mi == 1 ? S(1) : mi == 2 ? S(-1) : S(0);
    if (outlet) {
        a = mi == 1 ? -(sdm * h) : S(0);
        b = mi == 1 ? S(1) + sd0 * h : S(1);
        c = S(0);
        return;
    }
    const S tmp = sd0 * epd * hsq;
    const S dtc = (sd0 == S(0)) ? S(0) : S(1);
    const S ddw = dtc * sdm * h2, dde = dtc * sdp * h2;
    a = -ddw * lam - tmp;
    c = dde * lam - tmp;
    b = S(1) + S(2) * tmp;
}

// matrix of system m: systems 0 and 1 share eigenvalue 0
__host__ __device__ constexpr int mat_of(int m) { return m < 2 ? 0 : m - 1; }

// scale a projective pair or map by a power of two (exact) so that its
// largest entry lies in [0.5, 1)
template <typename S>
__device__ __forceinline__ int pow2_exponent(S mx) {
    int e = 0;
    (void)frexp(mx, &e);
    return e;
}

// One direction's DAF line solves on level L.
//   D = 0: rows i = 1 .. ni-2 (+ the outlet plane when ext) of every (j, k)
//          line; input S (-dtau J rh) transformed by Ma^-1, output to xout.
//   D = 1: rows j of every (i, k) line; input xin transformed by N1^-1.
//   D = 2: rows k of every (i, j) line; input xin transformed by N2^-1;
//          the solution times Mc is added to q (and stored in dq_out when
//          given, level-local, 4 SoA).
// use_b1 / use_b2: the semi-explicit boundary vectors brh of the first and
// last rows (fine level, rhs_diag_solver.F90:87-135 and its j/k copies),
// built here by the threads of those rows.
template <int D, int R, typename S>
__device__ __forceinline__ void daf_body(const Level& L, const Fields& F, const Consts& C, const SweepIO<S>& io,
                                         const SweepShape& sh, int use_b1, int use_b2, int b2_full, int ext,
                                         double* dq_out, unsigned bid) {
    extern __shared__ double smem_d[];
    S* lds = reinterpret_cast<S*>(smem_d);
    const int P = sh.P, TI = sh.TI, n = sh.n;
    const int LS = line_stride(P);
    const int tid = (int)threadIdx.x;
    // i lines are contiguous: segments of one line on adjacent lanes; j and k
    // lines: adjacent lines (adjacent i) on adjacent lanes
    const int li = D == 0 ? tid / P : tid % TI;
    const int ps = D == 0 ? tid % P : tid / TI;
    const long t = (long)bid * TI + li;
    const bool act = t < sh.nlines;
    const int ni = L.ni, nj = L.nj, nk = L.nk;
    long base = 0, sD;
    if (D == 0) {
        sD = 1;
        if (act) base = (long)ni * ((1 + t % (nj - 2)) + (long)nj * (1 + t / (nj - 2)));
    } else if (D == 1) {
        sD = ni;
        if (act) base = (1 + t % sh.nI) + (long)ni * nj * (1 + t / sh.nI);
    } else {
        sD = (long)ni * nj;
        if (act) base = (1 + t % sh.nI) + (long)ni * (1 + t / sh.nI);
    }
    const int nd = D == 0 ? ni : D == 1 ? nj : nk;  // nodes along the line
    // the characteristic-outlet row closes the i lines only
    const bool outlet_lines = D == 0 && ext;
    const int s0 = ps * R;
    const int nr = act ? max(0, min(R, n - s0)) : 0;
    const S h = (S)(D == 0 ? L.dc : D == 1 ? L.de : L.dz);
    const S h2 = S(0.5) * h, hsq = h * h, epd = (S)C.ep[0][D];
    auto gnode = [&](int p) { return L.off + base + (long)p * sD; };  // global node at position p

    // ---- load: right-hand sides (transformed) and dtau spr of rows s0-1 .. s0+nr ----
    // window entry w is node position s0 + w (row s0 + w - 1)
    S dv[R][4];
    S sdv[R + 2];
    if (D == 0) {
        // i lines: a segment's rows and its two window nodes are one run of
        // consecutive nodes, loaded as such (16-byte loads; lanes are R rows
        // apart).  The run of a partial last segment (or an idle thread)
        // reaches past the line's end into the next line or plane of the
        // level; those values are never used.
        double dt[R + 2], m1[R + 2], m2[R + 2], m3[R + 2], aj[R], rh[4][R];
#pragma unroll
        for (int w = 0; w < R + 2; w++) {
            const long g = gnode(s0 + w);
            dt[w] = F.dtau[g];
            m1[w] = F.csi[0][g];
            m2[w] = F.csi[1][g];
            m3[w] = F.csi[2][g];
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
            const long g = gnode(s0 + r + 1);
            aj[r] = F.aj[g];
#pragma unroll
            for (int m = 0; m < 4; m++) rh[m][r] = F.rh[m][g];
        }
        // sdw[w]: dtau spr at window entry w; rows r < nr transformed
        double sdw[R + 2];
#pragma unroll
        for (int w = 0; w < R + 2; w++) {
            if (w >= 1 && w <= R) {
                const double rv[4] = {rh[0][w - 1], rh[1][w - 1], rh[2][w - 1], rh[3][w - 1]};
                double v[4];
                sdw[w] = transform_vals0(C, dt[w], aj[w - 1], rv, m1[w], m2[w], m3[w], v);
                const bool in = w - 1 < nr;
#pragma unroll
                for (int m = 0; m < 4; m++) dv[w - 1][m] = in ? (S)v[m] : S(0);
            } else {
                sdw[w] = sd_vals(C, dt[w], m1[w], m2[w], m3[w]);
            }
        }
        const int pe = s0 + nr + 1;
        sdv[0] = nr > 0 ? (S)sdw[0] : S(0);
#pragma unroll
        for (int w = 1; w <= R + 1; w++) {
            // rows, then the node after the segment at entry nr + 1
            const bool row = w - 1 < nr;
            const bool after = w == nr + 1 && nr > 0 && pe <= nd - 1;
            sdv[w] = (row || after) ? (S)sdw[w] : S(0);
        }
    } else {
        // j and k lines: one branch per row (all their rows' loads in flight
        // would not fit the registers)
#pragma unroll
        for (int r = 0; r < R; r++) {
            sdv[r + 1] = S(0);
#pragma unroll
            for (int m = 0; m < 4; m++) dv[r][m] = S(0);
            if (r < nr) {
                const int p = s0 + r + 1;
                double v[4];
                const long tl = base + (long)p * sD;
#pragma unroll
                for (int m = 0; m < 4; m++) v[m] = (double)io.xin[m][tl];
                sdv[r + 1] = (S)transform_row<D>(F, C, gnode(p), v);
#pragma unroll
                for (int m = 0; m < 4; m++) dv[r][m] = (S)v[m];
            }
        }
        const int pe = s0 + nr + 1;
        sdv[0] = nr > 0 ? (S)row_sd<D>(F, C, gnode(s0)) : S(0);
        const S se = (nr > 0 && pe <= nd - 1) ? (S)row_sd<D>(F, C, gnode(pe)) : S(0);
        sdv[R + 1] = S(0);
#pragma unroll
        for (int w = 1; w <= R + 1; w++)
            if (w == nr + 1) sdv[w] = se;
    }
    // semi-explicit boundary vectors of the first and last rows
    const bool first = act && s0 == 0 && nr > 0;
    const bool last = act && nr > 0 && s0 + nr == n;
    if (first && use_b1) {
        double w[4];
        const int b0 = 2 * D;
        if (D == 0) {
            double r1[4], r2[4];
            const long g1 = gnode(1), g2 = gnode(2);
            scaled_rh(F, C, g1, node_dt(C, F.dtau[g1]), r1);
            scaled_rh(F, C, g2, node_dt(C, F.dtau[g2]), r2);
#pragma unroll
            for (int m = 0; m < 4; m++) w[m] = C.sa[b0][m] * r1[m] + C.sb[b0][m] * r2[m];
        } else {
#pragma unroll
            for (int m = 0; m < 4; m++)
                w[m] = C.sa[b0][m] * (double)io.xin[m][base + sD] + C.sb[b0][m] * (double)io.xin[m][base + 2 * sD];
        }
        pre_transform<D>(F, C, gnode(0), w);
#pragma unroll
        for (int m = 0; m < 4; m++) {
            S am, bm, cm;
            row_coef<S>(mat_of(m), sdv[0], sdv[1], sdv[2], h, h2, hsq, epd, outlet_lines && n == 1, am, bm, cm);
            dv[0][m] = dv[0][m] - am * (S)w[m];
        }
    }
    if (last && use_b2) {
        double w[4] = {0.0, 0.0, 0.0, 0.0};
        const int b1 = 2 * D + 1;
        if (D == 0) {
            // the fifth product of sa(1:5,1) * rh(1:4, i, ...) (:97) lands in
            // brh(1,2,j,k) when the i = 1 vector is built; nothing else
            // writes brh(:,2) on a single rank (:102, :135), so it is used
            // as it is
            if (use_b1) {
                const long g2 = gnode(2), g3 = gnode(3);
                const double x2 = -F.dtau[g2] * F.rh[0][g2] * F.aj[g2];
                const double x3 = -F.dtau[g3] * F.rh[0][g3] * F.aj[g3];
                w[0] = C.sa[0][4] * (C.beta * x2) + C.sb[0][4] * (C.beta * x3);
            }
        } else if (D == 1 || b2_full) {
#pragma unroll
            for (int m = 0; m < 4; m++)
                w[m] = C.sa[b1][m] * (double)io.xin[m][base + (long)(nd - 2) * sD] +
                       C.sb[b1][m] * (double)io.xin[m][base + (long)(nd - 3) * sD];
            pre_transform<D>(F, C, gnode(nd - 1), w);
        } else if (use_b1) {
            // brh(:,2) is not rebuilt when btype(6) == 6 (:467) and keeps
            // the fifth product of the k = 1 build (:462), which reads rh(1)
            // of node i+1: the j sweep's solution there, zero on a boundary
            // node (rh is zeroed on the boundary planes before the DAF)
            const bool on_line = (int)(t % sh.nI) + 2 <= sh.nI;  // i + 1 is a j-line position
            const double x1 = on_line ? (double)io.xin[0][base + sD + 1] : 0.0;
            const double x2 = on_line ? (double)io.xin[0][base + 2 * sD + 1] : 0.0;
            w[0] = C.sa[4][4] * x1 + C.sb[4][4] * x2;
            pre_transform<D>(F, C, gnode(nd - 1), w);
        }
        const int r = nr - 1;
#pragma unroll
        for (int rr = 0; rr < R; rr++) {
            if (rr == r) {
#pragma unroll
                for (int m = 0; m < 4; m++) {
                    S am, bm, cm;
                    row_coef<S>(mat_of(m), sdv[rr], sdv[rr + 1], sdv[rr + 2], h, h2, hsq, epd, outlet_lines, am, bm,
                                cm);
                    dv[rr][m] = dv[rr][m] - cm * (S)w[m];
                }
            }
        }
    }
    const bool outlet_row_here = outlet_lines && last;  // row n - 1 is the outlet row
    auto coef = [&](int r, int mi, S& a, S& b, S& c) {
        row_coef<S>(mi, sdv[r], sdv[r + 1], sdv[r + 2], h, h2, hsq, epd, outlet_row_here && r == nr - 1, a, b, c);
    };
    S* slot = lds + li * LS + ps * kSlot;
    const bool scanner = tid < 4 * TI;
    const int sl = tid / 4, sm = tid % 4;
    const bool scan_act = scanner && (long)bid * TI + sl < sh.nlines;
    S* sline = lds + sl * LS;  // the scanned line's slots

    // ---- phase A: Moebius maps of c' over the segment, scanned ----
    {
        S M[3][4];
#pragma unroll
        for (int mi = 0; mi < 3; mi++) { M[mi][0] = S(1); M[mi][1] = S(0); M[mi][2] = S(0); M[mi][3] = S(1); }
#pragma unroll
        for (int r = 0; r < R; r++) {
            if (r < nr) {
#pragma unroll
                for (int mi = 0; mi < 3; mi++) {
                    S a, b, c;
                    coef(r, mi, a, b, c);
                    // [c'; 1] -> [c; b - a c']
                    const S n0 = c * M[mi][2], n1 = c * M[mi][3];
                    const S n2 = b * M[mi][2] - a * M[mi][0], n3 = b * M[mi][3] - a * M[mi][1];
                    const int e = -pow2_exponent(fmax(fmax(fabs(n0), fabs(n1)), fmax(fabs(n2), fabs(n3))));
                    M[mi][0] = ldexp(n0, e); M[mi][1] = ldexp(n1, e); M[mi][2] = ldexp(n2, e); M[mi][3] = ldexp(n3, e);
                }
            }
        }
#pragma unroll
        for (int mi = 0; mi < 3; mi++)
#pragma unroll
            for (int e = 0; e < 4; e++) slot[4 * mi + e] = M[mi][e];
    }
    __syncthreads();
    if (scan_act && sm < 3) {
        // projective pair (cn, cd): c' entering the segment is cn / cd
        S cn = S(0), cd = S(1);
        for (int sg0 = 0; sg0 < P; sg0 += 4) {
            S q[4][4];
#pragma unroll
            for (int u = 0; u < 4; u++)
                if (sg0 + u < P)
#pragma unroll
                    for (int e = 0; e < 4; e++) q[u][e] = sline[(sg0 + u) * kSlot + 4 * sm + e];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                if (sg0 + u < P) {
                    sline[(sg0 + u) * kSlot + 4 * sm] = cn / cd;
                    const S nn = q[u][0] * cn + q[u][1] * cd, dd = q[u][2] * cn + q[u][3] * cd;
                    const int e = -pow2_exponent(fmax(fabs(nn), fabs(dd)));
                    cn = ldexp(nn, e);
                    cd = ldexp(dd, e);
                }
            }
        }
    }
    __syncthreads();
    S cst[3];
#pragma unroll
    for (int mi = 0; mi < 3; mi++) cst[mi] = slot[4 * mi];

    // ---- phase B: exact c' over the rows (their 1 / den kept); affine maps of d', scanned ----
    S inv[R][3];
    unsigned piv = 0;
    int zrow = 0;  // first row (1-based in the line) with a zero or non-finite pivot
    {
        S U[4], V[4];
#pragma unroll
        for (int m = 0; m < 4; m++) { U[m] = S(0); V[m] = S(1); }
#pragma unroll
        for (int mi = 0; mi < 3; mi++) {
            S cprev = cst[mi];
#pragma unroll
            for (int r = 0; r < R; r++) {
                inv[r][mi] = S(0);
                if (r < nr) {
                    S a, b, c;
                    coef(r, mi, a, b, c);
                    const S den = b - a * cprev;
                    const S iv = S(1) / den;
                    inv[r][mi] = iv;
                    cprev = c * iv;
                    if (!(den != S(0) && isfinite(den)) && zrow == 0) zrow = s0 + r + 1;
                    // SGTSV's interchange test against the next row's subdiagonal
                    // (a depends on the row's sd and its lower neighbour's only)
                    if (s0 + r + 1 < n) {
                        S an, bn, cn;
                        row_coef<S>(mi, sdv[r + 1], sdv[r + 2], S(0), h, h2, hsq, epd, outlet_lines && s0 + r + 2 == n,
                                    an, bn, cn);
                        if (fabs(den) < fabs(an)) piv += mi == 0 ? 2u : 1u;
                    }
#pragma unroll
                    for (int m = 0; m < 4; m++) {
                        if (mat_of(m) == mi) {
                            const S v = -a * iv;
                            U[m] = dv[r][m] * iv + v * U[m];
                            V[m] = v * V[m];
                        }
                    }
                }
            }
        }
#pragma unroll
        for (int m = 0; m < 4; m++) { slot[2 * m] = U[m]; slot[2 * m + 1] = V[m]; }
    }
    __syncthreads();
    if (scan_act) {
        S din = S(0);
        for (int sg0 = 0; sg0 < P; sg0 += 8) {
            S u[8], v[8];
#pragma unroll
            for (int k = 0; k < 8; k++)
                if (sg0 + k < P) { u[k] = sline[(sg0 + k) * kSlot + 2 * sm]; v[k] = sline[(sg0 + k) * kSlot + 2 * sm + 1]; }
#pragma unroll
            for (int k = 0; k < 8; k++) {
                if (sg0 + k < P) {
                    sline[(sg0 + k) * kSlot + 2 * sm] = din;
                    din = u[k] + v[k] * din;
                }
            }
        }
    }
    __syncthreads();

    // ---- phase C: exact d' over the rows; affine maps of x, scanned backwards ----
    {
        S dprev[4];
#pragma unroll
        for (int m = 0; m < 4; m++) dprev[m] = slot[2 * m];
#pragma unroll
        for (int r = 0; r < R; r++) {
            if (r < nr) {
#pragma unroll
                for (int mi = 0; mi < 3; mi++) {
                    S a, b, c;
                    coef(r, mi, a, b, c);
#pragma unroll
                    for (int m = 0; m < 4; m++) {
                        if (mat_of(m) == mi) {
                            dprev[m] = (dv[r][m] - a * dprev[m]) * inv[r][mi];
                            dv[r][m] = dprev[m];
                        }
                    }
                }
            }
        }
        S A[4], B[4];
#pragma unroll
        for (int m = 0; m < 4; m++) { A[m] = S(0); B[m] = S(1); }
#pragma unroll
        for (int r = R - 1; r >= 0; r--) {
            if (r < nr) {
#pragma unroll
                for (int mi = 0; mi < 3; mi++) {
                    S a, b, c;
                    coef(r, mi, a, b, c);
                    const S cc = c * inv[r][mi];
#pragma unroll
                    for (int m = 0; m < 4; m++) {
                        if (mat_of(m) == mi) {
                            A[m] = dv[r][m] - cc * A[m];
                            B[m] = -cc * B[m];
                        }
                    }
                }
            }
        }
#pragma unroll
        for (int m = 0; m < 4; m++) { slot[2 * m] = A[m]; slot[2 * m + 1] = B[m]; }
    }
    __syncthreads();
    if (scan_act) {
        S xin = S(0);
        for (int sg1 = P - 1; sg1 >= 0; sg1 -= 8) {
            S u[8], v[8];
#pragma unroll
            for (int k = 0; k < 8; k++)
                if (sg1 - k >= 0) { u[k] = sline[(sg1 - k) * kSlot + 2 * sm]; v[k] = sline[(sg1 - k) * kSlot + 2 * sm + 1]; }
#pragma unroll
            for (int k = 0; k < 8; k++) {
                if (sg1 - k >= 0) {
                    sline[(sg1 - k) * kSlot + 2 * sm] = xin;
                    xin = u[k] + v[k] * xin;
                }
            }
        }
    }
    __syncthreads();

    // ---- phase D: back substitution over the rows and the output ----
    S x[4];
#pragma unroll
    for (int m = 0; m < 4; m++) x[m] = slot[2 * m];
#pragma unroll
    for (int r = R - 1; r >= 0; r--) {
        if (r < nr) {
#pragma unroll
            for (int mi = 0; mi < 3; mi++) {
                S a, b, c;
                coef(r, mi, a, b, c);
                const S cc = c * inv[r][mi];
#pragma unroll
                for (int m = 0; m < 4; m++) {
                    if (mat_of(m) == mi) {
                        x[m] = dv[r][m] - cc * x[m];
                        dv[r][m] = x[m];
                    }
                }
            }
            if (D == 1 || (D == 0 && nr < R)) {
                const long tl = base + (long)(s0 + r + 1) * sD;
#pragma unroll
                for (int m = 0; m < 4; m++) io.xout[m][tl] = dv[r][m];
            }
        }
    }
    if (D == 0 && nr == R) {
        // a whole segment of an i line: one run of consecutive nodes per
        // component
#pragma unroll
        for (int m = 0; m < 4; m++)
#pragma unroll
            for (int r = 0; r < R; r++) io.xout[m][base + s0 + r + 1] = dv[r][m];
    }
    if (D == 2) {
        // Mc and the q update, four rows at a time: the loads of a group are
        // all in flight, and every value formed, before its first store (rows
        // past the line's end read its last row and are not stored)
#pragma unroll
        for (int r0 = 0; r0 < R; r0 += 4) {
            double v[4][4], qv[4][4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const long g = L.off + base + (long)(min(s0 + r0 + u, n - 1) + 1) * sD;
                const double dt = F.dtau[g], z1 = F.zet[0][g], z2 = F.zet[1][g], z3 = F.zet[2][g];
#pragma unroll
                for (int m = 0; m < 4; m++) qv[u][m] = F.q[m][g];
#pragma unroll
                for (int m = 0; m < 4; m++) v[u][m] = (double)dv[r0 + u][m];
                apply_mc(C, dt, z1, z2, z3, v[u]);
#pragma unroll
                for (int m = 0; m < 4; m++) qv[u][m] = qv[u][m] + v[u][m];
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int r = r0 + u;
                if (r < nr) {
                    const long tl = base + (long)(s0 + r + 1) * sD;
#pragma unroll
                    for (int m = 0; m < 4; m++) F.q[m][L.off + tl] = qv[u][m];
                    if (dq_out) {
#pragma unroll
                        for (int m = 0; m < 4; m++) dq_out[m * L.n + tl] = v[u][m];
                    }
                }
            }
        }
    }
    if (piv) atomicAdd(F.pivots, (unsigned long long)piv);
    if (zrow) atomicMin(F.zero_pivot, (unsigned)zrow);
}

// the boundary-vector flags and the line shape of direction D on level n
// (the i lines take the characteristic-outlet row when ext); P and TI are
// the launch's
struct SweepFlags {
    int use_b1, use_b2, b2_full, ext;
};
__host__ __device__ inline SweepFlags sweep_flags(const Consts& C, int D, int n) {
    SweepFlags f;
    f.use_b1 = (n == 1 && C.btype[2 * D] != 5);
    f.use_b2 = (n == 1 && C.btype[2 * D + 1] != 5);
    f.b2_full = (n == 1 && C.btype[5] != 6);
    // a characteristic outlet at i = ni - 1 joins the fine level's sweeps
    f.ext = (n == 1 && C.btype[1] == DESALM_BC_OUTLET_MOC);
    return f;
}
__host__ __device__ inline int sweep_len(const Level& L, int D, int ext) {
    return D == 0 ? L.ni - 2 + ext : D == 1 ? L.nj - 2 : L.nk - 2;
}
__host__ __device__ inline SweepShape sweep_lines(const Level& L, int D, int ext, int P, int TI) {
    SweepShape sh;
    sh.nI = L.ni - 2 + ext;
    sh.n = sweep_len(L, D, ext);
    sh.nlines = D == 0 ? (long)(L.nj - 2) * (L.nk - 2) : D == 1 ? (long)sh.nI * (L.nk - 2) : (long)sh.nI * (L.nj - 2);
    sh.P = P;
    sh.TI = TI;
    return sh;
}
// segments per line and lines per workgroup for lines of up to len rows
__host__ __device__ inline void sweep_tiles(int D, int len, int R, int& P, int& TI) {
    P = (len + R - 1) / R;
    if (P < 4) P = 4;
    // j and k lines: 16 adjacent i lines per workgroup (128-byte rows; the
    // partial cache lines at a run's ends are shared with the neighbouring
    // workgroups, which run on the same XCD).  A 256-thread workgroup leaves
    // room for a second one on the CU at up to 256 VGPRs.
    TI = D == 0 ? (256 / P > 1 ? 256 / P : 1) : 16;
    while (TI * P > 512 && TI > 1) TI /= 2;
}

// the solution of the i sweep goes to F.dp, of the j sweep to F.cp
// (level-local); the mixed form uses their FP32 counterparts
template <typename S>
__host__ __device__ inline SweepIO<S> sweep_io(const Fields& F, int D);
template <>
__host__ __device__ inline SweepIO<double> sweep_io<double>(const Fields& F, int D) {
    SweepIO<double> io;
    for (int m = 0; m < 4; m++) {
        io.xin[m] = D == 1 ? F.dp[m] : F.cp[m];
        io.xout[m] = D == 0 ? F.dp[m] : F.cp[m];
    }
    return io;
}
template <>
__host__ __device__ inline SweepIO<float> sweep_io<float>(const Fields& F, int D) {
    SweepIO<float> io;
    for (int m = 0; m < 4; m++) {
        io.xin[m] = D == 1 ? F.dpf[m] : F.cpf[m];
        io.xout[m] = D == 0 ? F.dpf[m] : F.cpf[m];
    }
    return io;
}

template <int D, int R, typename S>
__global__ void __launch_bounds__(512) k_daf_sweep(Level L, Fields F, Consts C, SweepIO<S> io, SweepShape sh,
                                                  int use_b1, int use_b2, int b2_full, int ext, double* dq_out) {
    daf_body<D, R, S>(L, F, C, io, sh, use_b1, use_b2, b2_full, ext, dq_out, xcd_block(blockIdx.x, gridDim.x));
}

// every zone of a group in one launch; P and TI are the launch's (from the
// longest line of any zone): a zone's shorter lines leave trailing segments
// empty, which pass the scanned maps through unchanged
template <int D, int R, typename S>
__global__ void __launch_bounds__(512) k_daf_sweep_z(ZoneGrid G, int n, int P, int TI) {
    const ZoneBlock b = zone_block(G);
    const ZoneDesc& Z = G.z[b.z];
    const Level& L = Z.lv[n];
    const SweepFlags f = sweep_flags(Z.C, D, n);
    daf_body<D, R, S>(L, Z.F, Z.C, sweep_io<S>(Z.F, D), sweep_lines(L, D, f.ext, P, TI), f.use_b1, f.use_b2,
                      f.b2_full, f.ext, nullptr, xcd_block(b.bid, b.nblk));
}

// ---- batched LAPACK SGTSV (src/sgtsv.f:89-259) in FP64, one system per thread ----
__global__ void k_gtsv_batch(int n, int nsys, double* dl, double* d, double* du, double* b, int* info) {
    const int sidx = blockIdx.x * blockDim.x + threadIdx.x;
    if (sidx >= nsys) return;
    double* DL = dl + (long)sidx * n;
    double* D = d + (long)sidx * n;
    double* DU = du + (long)sidx * n;
    double* B = b + (long)sidx * n;
    int inf = 0;
    for (int i = 0; i < n - 2 && !inf; i++) {
        if (fabs(D[i]) >= fabs(DL[i])) {
            if (D[i] != 0.0) {
                const double fact = DL[i] / D[i];
                D[i + 1] = D[i + 1] - fact * DU[i];
                B[i + 1] = B[i + 1] - fact * B[i];
            } else {
                inf = i + 1;
                break;
            }
            DL[i] = 0.0;
        } else {
            const double fact = D[i] / DL[i];
            D[i] = DL[i];
            double temp = D[i + 1];
            D[i + 1] = DU[i] - fact * temp;
            DL[i] = DU[i + 1];
            DU[i + 1] = -fact * DL[i];
            DU[i] = temp;
            temp = B[i];
            B[i] = B[i + 1];
            B[i + 1] = temp - fact * B[i + 1];
        }
    }
    if (!inf && n > 1) {
        const int i = n - 2;
        if (fabs(D[i]) >= fabs(DL[i])) {
            if (D[i] != 0.0) {
                const double fact = DL[i] / D[i];
                D[i + 1] = D[i + 1] - fact * DU[i];
                B[i + 1] = B[i + 1] - fact * B[i];
            } else {
                inf = i + 1;
            }
        } else {
            const double fact = D[i] / DL[i];
            D[i] = DL[i];
            const double temp = D[i + 1];
            D[i + 1] = DU[i] - fact * temp;
            DU[i] = temp;
            const double tb = B[i];
            B[i] = B[i + 1];
            B[i + 1] = tb - fact * B[i + 1];
        }
    }
    if (!inf && n > 0 && D[n - 1] == 0.0) inf = n;
    if (!inf && n > 0) {
        B[n - 1] = B[n - 1] / D[n - 1];
        if (n > 1) B[n - 2] = (B[n - 2] - DU[n - 2] * B[n - 1]) / D[n - 2];
        for (int i = n - 3; i >= 0; i--) B[i] = (B[i] - DU[i] * B[i + 1] - DL[i] * B[i + 2]) / D[i];
    }
    if (info) info[sidx] = inf;
}


}  // namespace

template <int D, typename S>
static void daf_sweep(hipStream_t s, const Level& L, const Fields& F, const Consts& C, int n, double* dq_out) {
    const SweepFlags f = sweep_flags(C, D, n);
    const int len = sweep_len(L, D, f.ext);
    // i lines: short segments (their rows are adjacent in memory, so lanes
    // reading row r of adjacent segments are R rows apart)
    const int R = (D == 0 || len <= 64) ? 4 : 8;
    int P = 0, TI = 0;
    sweep_tiles(D, len, R, P, TI);
    const SweepShape sh = sweep_lines(L, D, f.ext, P, TI);
    const unsigned nb = (unsigned)((sh.nlines + sh.TI - 1) / sh.TI);
    const size_t lds = sizeof(S) * (size_t)sh.TI * line_stride(sh.P);
    const SweepIO<S> io = sweep_io<S>(F, D);
    if (R == 4)
        hipLaunchKernelGGL((k_daf_sweep<D, 4, S>), dim3(nb), dim3(sh.TI * sh.P), lds, s, L, F, C, io, sh, f.use_b1,
                           f.use_b2, f.b2_full, f.ext, dq_out);
    else
        hipLaunchKernelGGL((k_daf_sweep<D, 8, S>), dim3(nb), dim3(sh.TI * sh.P), lds, s, L, F, C, io, sh, f.use_b1,
                           f.use_b2, f.b2_full, f.ext, dq_out);
}

template <int D, typename S>
static void daf_sweep_z(hipStream_t s, const ZoneList& Z, int n) {
    int len = 0;
    for (const ZoneDesc& z : Z.host) len = std::max(len, sweep_len(z.lv[n], D, sweep_flags(z.C, D, n).ext));
    const int R = (D == 0 || len <= 64) ? 4 : 8;
    int P = 0, TI = 0;
    sweep_tiles(D, len, R, P, TI);
    unsigned nb = 0;
    const ZoneGrid G = Z.grid(
        [&](const ZoneDesc& z) {
            const SweepShape sh = sweep_lines(z.lv[n], D, sweep_flags(z.C, D, n).ext, P, TI);
            return (unsigned)((sh.nlines + TI - 1) / TI);
        },
        nb);
    const size_t lds = sizeof(S) * (size_t)TI * line_stride(P);
    if (R == 4)
        hipLaunchKernelGGL((k_daf_sweep_z<D, 4, S>), dim3(nb), dim3(TI * P), lds, s, G, n, P, TI);
    else
        hipLaunchKernelGGL((k_daf_sweep_z<D, 8, S>), dim3(nb), dim3(TI * P), lds, s, G, n, P, TI);
}

template <int D>
static void daf_dir(hipStream_t s, const Level& L, const Fields& F, const Consts& C, int n, double* dq_out) {
    if (C.mixed) daf_sweep<D, float>(s, L, F, C, n, dq_out);
    else daf_sweep<D, double>(s, L, F, C, n, dq_out);
}

// the sweep of direction D on level n over every zone of a group in one
// launch (the zones share the precision form)
void launch_daf_z(hipStream_t s, const ZoneList& Z, int n, int D) {
    const bool f = Z.host[0].C.mixed;
    if (D == 0) f ? daf_sweep_z<0, float>(s, Z, n) : daf_sweep_z<0, double>(s, Z, n);
    else if (D == 1) f ? daf_sweep_z<1, float>(s, Z, n) : daf_sweep_z<1, double>(s, Z, n);
    else f ? daf_sweep_z<2, float>(s, Z, n) : daf_sweep_z<2, double>(s, Z, n);
}

void launch_daf_i(hipStream_t s, const Level& L, const Fields& F, const Consts& C, int n) {
    daf_dir<0>(s, L, F, C, n, nullptr);
}

void launch_daf_j(hipStream_t s, const Level& L, const Fields& F, const Consts& C, int n) {
    daf_dir<1>(s, L, F, C, n, nullptr);
}

void launch_daf_k(hipStream_t s, const Level& L, const Fields& F, const Consts& C, int n, double* dq_out) {
    daf_dir<2>(s, L, F, C, n, dq_out);
}

void launch_gtsv_batch(hipStream_t s, int n, int nsys, double* dl, double* d, double* du, double* b, int* info) {
    const int bs = 64;
    hipLaunchKernelGGL(k_gtsv_batch, dim3((unsigned)((nsys + bs - 1) / bs)), dim3(bs), 0, s, n, nsys, dl, d, du, b,
                       info);
}

}  // namespace desalm
