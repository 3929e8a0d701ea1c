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

// ---- modal matrices (rhs_modal_matrices.F90:112-264) ----
// Ma-type matrix for metric row (m1, m2, m3)
__device__ __forceinline__ void build_m(double m1, double m2, double m3, double sigma, double M[4][4]) {
    const double g = m1 * m1 + m2 * m2 + m3 * m3;
    const double sg = sqrt(g);
    const double s = sqrt(2.0 * g + (m2 - m3) * (m2 - m3));
    M[0][0] = 0.0; M[0][1] = 0.0; M[0][2] = sigma; M[0][3] = sigma;
    M[1][0] = -(m2 + m3) / s; M[1][1] = m1 * (m3 - m2) / (s * sg); M[1][2] = m1 / sg; M[1][3] = -m1 / sg;
    M[2][0] = m1 / s; M[2][1] = (g + m2 * (m3 - m2)) / (s * sg); M[2][2] = m2 / sg; M[2][3] = -m2 / sg;
    M[3][0] = m1 / s; M[3][1] = -(g - m3 * (m3 - m2)) / (s * sg); M[3][2] = m3 / sg; M[3][3] = -m3 / sg;
}

// Ma^-1-type matrix for metric row (m1, m2, m3)
__device__ __forceinline__ void build_minv(double m1, double m2, double m3, double sigma, double Mi[4][4]) {
    const double m1s = m1 * m1, m2s = m2 * m2, m3s = m3 * m3;
    const double g = m1s + m2s + m3s;
    const double sg = sqrt(g);
    const double s = sqrt(2.0 * g + (m2 - m3) * (m2 - m3));
    const double d = g + 2.0 * m2 * m3 + m1 * m1;
    const double r1 = m2 + m3, r2 = m2 - m3, r3 = m2s - m3s;
    Mi[0][0] = 0.0; Mi[0][1] = -s * (m2 + m3) / d; Mi[0][2] = s * m1 / d; Mi[0][3] = Mi[0][2];
    const double ssgd = s / sg / d;
    const double sgd2 = 1.0 / sg / d / 2.0;
    Mi[1][0] = 0.0; Mi[1][1] = m1 * (m3 - m2) * ssgd;
    Mi[1][2] = (m2 * m3 + m1s + m3s) * ssgd; Mi[1][3] = -(m2 * m3 + m1s + m2s) * ssgd;
    Mi[2][0] = 0.5 / sigma; Mi[2][1] = m1 * (2.0 * g + 2.0 * m2 * m3 - m2s - m3s) * sgd2;
    Mi[2][2] = (g * r1 + m1s * r2 + m3 * r3) * sgd2; Mi[2][3] = (g * r1 - m1s * r2 - m2 * r3) * sgd2;
    Mi[3][0] = 0.5 / sigma; Mi[3][1] = -Mi[2][1]; Mi[3][2] = -Mi[2][2]; Mi[3][3] = -Mi[2][3];
}

__device__ __forceinline__ void matmul4(const double a[4][4], const double b[4][4], double c[4][4]) {
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int s = 0; s < 4; s++) {
            double acc = 0.0;
#pragma unroll
            for (int t = 0; t < 4; t++) acc = acc + a[r][t] * b[t][s];
            c[r][s] = acc;
        }
}

__device__ __forceinline__ void matvec4(const double a[4][4], double v[4]) {
    double rp[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        double acc = 0.0;
#pragma unroll
        for (int c = 0; c < 4; c++) acc = acc + a[r][c] * v[c];
        rp[r] = acc;
    }
#pragma unroll
    for (int r = 0; r < 4; r++) v[r] = rp[r];
}

// sigma of the modal matrices at a node (rhs_modal_matrices.F90:82-84)
__device__ __forceinline__ double modal_sigma(const Consts& C, double dtau) {
    const double alpha = 1.0 + C.e_source * 1.5 * dtau / C.delti;
    return sqrt(C.beta * alpha);
}
__device__ __forceinline__ double modal_theta(const Consts& C, double dtau) {
    const double alpha = 1.0 + C.e_source * 1.5 * dtau / C.delti;
    return sqrt(C.beta / alpha);
}

// v <- Ma^-1 v at global node g
__device__ __forceinline__ void apply_mai(const Fields& F, const Consts& C, long g, double v[4]) {
    double M[4][4];
    build_minv(F.csi[0][g], F.csi[1][g], F.csi[2][g], modal_sigma(C, F.dtau[g]), M);
    matvec4(M, v);
}
// v <- (Mb^-1 Ma) v
__device__ __forceinline__ void apply_n1i(const Fields& F, const Consts& C, long g, double v[4]) {
    const double sigma = modal_sigma(C, F.dtau[g]);
    double A[4][4], B[4][4], P[4][4];
    build_m(F.csi[0][g], F.csi[1][g], F.csi[2][g], sigma, A);
    build_minv(F.eta[0][g], F.eta[1][g], F.eta[2][g], sigma, B);
    matmul4(B, A, P);
    matvec4(P, v);
}
// v <- (Mc^-1 Mb) v
__device__ __forceinline__ void apply_n2i(const Fields& F, const Consts& C, long g, double v[4]) {
    const double sigma = modal_sigma(C, F.dtau[g]);
    double A[4][4], B[4][4], P[4][4];
    build_m(F.eta[0][g], F.eta[1][g], F.eta[2][g], sigma, A);
    build_minv(F.zet[0][g], F.zet[1][g], F.zet[2][g], sigma, B);
    matmul4(B, A, P);
    matvec4(P, v);
}
// v <- Mc v
__device__ __forceinline__ void apply_mc(const Fields& F, const Consts& C, long g, double v[4]) {
    double M[4][4];
    build_m(F.zet[0][g], F.zet[1][g], F.zet[2][g], modal_sigma(C, F.dtau[g]), M);
    matvec4(M, v);
}

// spectral radius component D at a node times dtau: returns (dtau, spr)

// rows per segment: registers per thread hold R rows of the four systems
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

// values of one segment slot in LDS (the largest phase: 3 Moebius maps)
constexpr int kSlot = 12;

// sigma, theta of the modal matrices at a node from its dtau
// (rhs_modal_matrices.F90:82-84)
__device__ __forceinline__ double node_alpha(const Consts& C, double dtau) {
    return 1.0 + C.e_source * 1.5 * dtau / C.delti;
}

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
__device__ __forceinline__ int mat_of(int m) { return m < 2 ? 0 : m - 1; }

// S (-dtau J rh) at global node g (rhs_diag_solver.F90:56-72)
__device__ __forceinline__ void scaled_rh(const Fields& F, const Consts& C, long g, double v[4]) {
    const double dt = F.dtau[g];
    const double alpha = 1.0 + C.e_source * 1.5 / C.delti * dt;
    double r[4];
#pragma unroll
    for (int m = 0; m < 4; m++) r[m] = -dt * F.rh[m][g] * F.aj[g];
    v[0] = C.beta * r[0];
#pragma unroll
    for (int m = 1; m < 4; m++) v[m] = 1.0 / alpha * r[m];
}

// the modal transform of sweep D applied to v at global node g
template <int D>
__device__ __forceinline__ void pre_transform(const Fields& F, const Consts& C, long g, double v[4]) {
    if (D == 0) apply_mai(F, C, g, v);
    else if (D == 1) apply_n1i(F, C, g, v);
    else apply_n2i(F, C, g, v);
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
__global__ void __launch_bounds__(512) k_daf_sweep(Level L, Fields F, Consts C, SweepIO<S> io, SweepShape sh,
                                                  int use_b1, int use_b2, int b2_full, int ext, double* dq_out) {
    extern __shared__ double smem_d[];
    S* lds = reinterpret_cast<S*>(smem_d);
    const int P = sh.P, TI = sh.TI, n = sh.n;
    const int tid = (int)threadIdx.x;
    // i lines are contiguous: segments of one line on adjacent lanes; j and k
    // lines: adjacent lines (adjacent i) on adjacent lanes
    const int li = D == 0 ? tid / P : tid % TI;
    const int ps = D == 0 ? tid % P : tid / TI;
    const long t = (long)blockIdx.x * TI + li;
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
    S dv[R][4];
    S sdv[R + 2];
#pragma unroll
    for (int w = 0; w < R + 2; w++) {
        const int p = s0 + w;  // node position of window entry w (row s0 + w - 1)
        sdv[w] = S(0);
        if (w <= nr + 1 && p <= nd - 1 && nr > 0) {
            const long g = gnode(p);
            const double* m0 = D == 0 ? F.csi[0] : D == 1 ? F.eta[0] : F.zet[0];
            const double* m1 = D == 0 ? F.csi[1] : D == 1 ? F.eta[1] : F.zet[1];
            const double* m2 = D == 0 ? F.csi[2] : D == 1 ? F.eta[2] : F.zet[2];
            const double dt = F.dtau[g];
            const double gg = m0[g] * m0[g] + m1[g] * m1[g] + m2[g] * m2[g];
            sdv[w] = (S)(dt * (sqrt(C.beta / node_alpha(C, dt)) * sqrt(gg)));
        }
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
        if (r < nr) {
            const int p = s0 + r + 1;
            const long g = gnode(p);
            double v[4];
            if (D == 0) {
                scaled_rh(F, C, g, v);
            } else {
                const long tl = base + (long)p * sD;
#pragma unroll
                for (int m = 0; m < 4; m++) v[m] = (double)io.xin[m][tl];
            }
            pre_transform<D>(F, C, g, v);
#pragma unroll
            for (int m = 0; m < 4; m++) dv[r][m] = (S)v[m];
        } else {
#pragma unroll
            for (int m = 0; m < 4; m++) dv[r][m] = S(0);
        }
    }
    // semi-explicit boundary vectors of the first and last rows
    const bool first = act && s0 == 0 && nr > 0;
    const bool last = act && nr > 0 && s0 + nr == n;
    if (first && use_b1) {
        double w[4];
        const int b0 = 2 * D;
        if (D == 0) {
            double r1[4], r2[4];
            scaled_rh(F, C, gnode(1), r1);
            scaled_rh(F, C, gnode(2), r2);
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
                double r2[4], r3[4];
                scaled_rh(F, C, gnode(2), r2);
                scaled_rh(F, C, gnode(3), r3);
                w[0] = C.sa[0][4] * r2[0] + C.sb[0][4] * r3[0];
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
        for (int m = 0; m < 4; m++) {
            S am, bm, cm;
            row_coef<S>(mat_of(m), sdv[r], sdv[r + 1], sdv[r + 2], h, h2, hsq, epd, outlet_lines, am, bm, cm);
            dv[r][m] = dv[r][m] - cm * (S)w[m];
        }
    }
    const bool outlet_row_here = outlet_lines && last;  // row n - 1 is the outlet row
    auto coef = [&](int r, int mi, S& a, S& b, S& c) {
        row_coef<S>(mi, sdv[r], sdv[r + 1], sdv[r + 2], h, h2, hsq, epd, outlet_row_here && r == nr - 1, a, b, c);
    };
    S* slot = lds + (long)(li * P + ps) * kSlot;
    const bool scanner = tid < 4 * TI;
    const int sl = tid / 4, sm = tid % 4;
    const bool scan_act = scanner && (long)blockIdx.x * TI + sl < sh.nlines;

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
                    const S sc = fmax(fmax(fabs(n0), fabs(n1)), fmax(fabs(n2), fabs(n3)));
                    const S is = sc > S(0) ? S(1) / sc : S(1);
                    M[mi][0] = n0 * is; M[mi][1] = n1 * is; M[mi][2] = n2 * is; M[mi][3] = n3 * is;
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
        S cin = S(0);
        for (int sg = 0; sg < P; sg++) {
            S* q = lds + (long)(sl * P + sg) * kSlot + 4 * sm;
            const S m0 = q[0], m1 = q[1], m2 = q[2], m3 = q[3];
            q[0] = cin;
            cin = (m0 * cin + m1) / (m2 * cin + m3);
        }
    }
    __syncthreads();
    S cst[3];
#pragma unroll
    for (int mi = 0; mi < 3; mi++) cst[mi] = slot[4 * mi];

    // ---- phase B: exact c' over the rows; affine maps of d', scanned ----
    S cp[R][3];
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
                if (r < nr) {
                    S a, b, c;
                    coef(r, mi, a, b, c);
                    const S den = b - a * cprev;
                    const S inv = S(1) / den;
                    cprev = c * inv;
                    cp[r][mi] = cprev;
                    if (!(den != S(0) && isfinite(den)) && zrow == 0) zrow = s0 + r + 1;
                    // SGTSV's interchange test against the next row's subdiagonal
                    if (s0 + r + 1 < n) {
                        S an, bn, cn;
                        row_coef<S>(mi, sdv[r + 1], sdv[r + 2], r + 3 < R + 2 ? sdv[min(r + 3, R + 1)] : S(0), h,
                                    h2, hsq, epd, outlet_lines && s0 + r + 2 == n, an, bn, cn);
                        if (fabs(den) < fabs(an)) piv += mi == 0 ? 2u : 1u;
                    }
#pragma unroll
                    for (int m = 0; m < 4; m++) {
                        if (mat_of(m) == mi) {
                            const S v = -a * inv;
                            U[m] = dv[r][m] * inv + v * U[m];
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
        for (int sg = 0; sg < P; sg++) {
            S* q = lds + (long)(sl * P + sg) * kSlot + 2 * sm;
            const S u = q[0], v = q[1];
            q[0] = din;
            din = u + v * din;
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
                    const S den = b - a * (r == 0 ? cst[mi] : cp[r - 1][mi]);
                    const S inv = S(1) / den;
#pragma unroll
                    for (int m = 0; m < 4; m++) {
                        if (mat_of(m) == mi) {
                            dprev[m] = (dv[r][m] - a * dprev[m]) * inv;
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
                for (int m = 0; m < 4; m++) {
                    const S cc = cp[r][mat_of(m)];
                    A[m] = dv[r][m] - cc * A[m];
                    B[m] = -cc * B[m];
                }
            }
        }
#pragma unroll
        for (int m = 0; m < 4; m++) { slot[2 * m] = A[m]; slot[2 * m + 1] = B[m]; }
    }
    __syncthreads();
    if (scan_act) {
        S xin = S(0);
        for (int sg = P - 1; sg >= 0; sg--) {
            S* q = lds + (long)(sl * P + sg) * kSlot + 2 * sm;
            const S a = q[0], b = q[1];
            q[0] = xin;
            xin = a + b * xin;
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
            const int p = s0 + r + 1;
#pragma unroll
            for (int m = 0; m < 4; m++) {
                x[m] = dv[r][m] - cp[r][mat_of(m)] * x[m];
                dv[r][m] = x[m];
            }
            const long tl = base + (long)p * sD;
            if (D < 2) {
#pragma unroll
                for (int m = 0; m < 4; m++) io.xout[m][tl] = dv[r][m];
            } else {
                const long g = L.off + tl;
                double v[4];
#pragma unroll
                for (int m = 0; m < 4; m++) v[m] = (double)dv[r][m];
                apply_mc(F, C, g, v);
#pragma unroll
                for (int m = 0; m < 4; m++) F.q[m][g] = F.q[m][g] + v[m];
                if (dq_out) {
#pragma unroll
                    for (int m = 0; m < 4; m++) dq_out[m * L.n + tl] = v[m];
                }
            }
        }
    }
    if (piv) atomicAdd(F.pivots, (unsigned long long)piv);
    if (zrow) atomicMin(F.zero_pivot, (unsigned)zrow);
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

// launch shape of one direction's sweep on a level
static SweepShape sweep_shape(const Level& L, int D, int ext, int R) {
    SweepShape sh;
    sh.nI = L.ni - 2 + ext;
    sh.n = D == 0 ? L.ni - 2 + ext : D == 1 ? L.nj - 2 : L.nk - 2;
    sh.nlines = D == 0 ? (long)(L.nj - 2) * (L.nk - 2) : D == 1 ? (long)sh.nI * (L.nk - 2) : (long)sh.nI * (L.nj - 2);
    sh.P = std::max(4, (sh.n + R - 1) / R);
    // j and k lines: 16 adjacent i lines per workgroup (128-byte rows)
    sh.TI = D == 0 ? std::max(1, 256 / sh.P) : 16;
    while (sh.TI * sh.P > 512 && sh.TI > 1) sh.TI /= 2;
    return sh;
}

template <int D, typename S>
static void daf_sweep(hipStream_t s, const Level& L, const Fields& F, const Consts& C, const SweepIO<S>& io, int n,
                      double* dq_out) {
    const int use_b1 = (n == 1 && C.btype[2 * D] != 5);
    const int use_b2 = (n == 1 && C.btype[2 * D + 1] != 5);
    const int b2_full = (n == 1 && C.btype[5] != 6);
    // a characteristic outlet at i = ni - 1 joins the fine level's sweeps
    const int ext = (n == 1 && C.btype[1] == DESALM_BC_OUTLET_MOC);
    const int len = D == 0 ? L.ni - 2 + ext : D == 1 ? L.nj - 2 : L.nk - 2;
    const int R = len <= 64 ? 4 : 8;
    const SweepShape sh = sweep_shape(L, D, ext, R);
    const unsigned nb = (unsigned)((sh.nlines + sh.TI - 1) / sh.TI);
    const size_t lds = sizeof(S) * kSlot * (size_t)sh.TI * sh.P;
    if (R == 4)
        hipLaunchKernelGGL((k_daf_sweep<D, 4, S>), dim3(nb), dim3(sh.TI * sh.P), lds, s, L, F, C, io, sh, use_b1,
                           use_b2, b2_full, ext, dq_out);
    else
        hipLaunchKernelGGL((k_daf_sweep<D, 8, S>), dim3(nb), dim3(sh.TI * sh.P), lds, s, L, F, C, io, sh, use_b1,
                           use_b2, b2_full, ext, dq_out);
}

// the solution of the i sweep goes to F.dp, of the j sweep to F.cp
// (level-local); the mixed form uses their FP32 counterparts
template <typename S>
static SweepIO<S> sweep_io(const Fields& F, int D);
template <>
SweepIO<double> sweep_io<double>(const Fields& F, int D) {
    SweepIO<double> io;
    for (int m = 0; m < 4; m++) {
        io.xin[m] = D == 1 ? F.dp[m] : F.cp[m];
        io.xout[m] = D == 0 ? F.dp[m] : F.cp[m];
    }
    return io;
}
template <>
SweepIO<float> sweep_io<float>(const Fields& F, int D) {
    SweepIO<float> io;
    for (int m = 0; m < 4; m++) {
        io.xin[m] = D == 1 ? F.dpf[m] : F.cpf[m];
        io.xout[m] = D == 0 ? F.dpf[m] : F.cpf[m];
    }
    return io;
}

template <int D>
static void daf_dir(hipStream_t s, const Level& L, const Fields& F, const Consts& C, int n, double* dq_out) {
    if (C.mixed) daf_sweep<D, float>(s, L, F, C, sweep_io<float>(F, D), n, dq_out);
    else daf_sweep<D, double>(s, L, F, C, sweep_io<double>(F, D), n, dq_out);
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
