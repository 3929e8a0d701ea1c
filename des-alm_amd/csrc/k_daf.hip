// k_daf.hip -- diagonalized approximate-factorization implicit operator.
//
// rhs_modal_matrices + rhs_diag_solver (src/rhs_modal_matrices.F90,
// src/rhs_diag_solver.F90).  Each direction D is two kernels:
//   k_daf_pre<D>  pointwise: the modal transform that precedes the sweep
//                 (S and Ma^-1 for i, N1^-1 = Mb^-1 Ma for j, N2^-1 = Mc^-1 Mb
//                 for k), dtau*spr_D per node, and the boundary vectors brh
//   k_daf_tri<D>  one thread per (line, scalar system): Thomas elimination
// and k_daf_post applies Mc and adds dq to q.  The eigenvector matrices are
// rebuilt in registers from the metrics and dtau of each node instead of
// being stored (the reference keeps 67 words per node).  LAPACK SGTSV, which
// the reference calls, picks the same pivots as Thomas unless |d_i| <
// |dl_i|; rows where that happens are counted in F.pivots so a caller can see
// when the pivot-free elimination departed from SGTSV's row order.
#include "desalm_internal.h"
#include "line_tile.h"

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
template <int D>
__device__ __forceinline__ double spr_at(const Fields& F, const Consts& C, long g) {
    const double* m0 = D == 0 ? F.csi[0] : D == 1 ? F.eta[0] : F.zet[0];
    const double* m1 = D == 0 ? F.csi[1] : D == 1 ? F.eta[1] : F.zet[1];
    const double* m2 = D == 0 ? F.csi[2] : D == 1 ? F.eta[2] : F.zet[2];
    const double gg = m0[g] * m0[g] + m1[g] * m1[g] + m2[g] * m2[g];
    return modal_theta(C, F.dtau[g]) * sqrt(gg);
}

// Scratch of one DAF pass in storage type S (level-local indices): S =
// double uses the FP64 scratch and leaves each sweep's solution in the rh
// field of the level; S = float is the mixed-precision form, whose sweeps
// run in FP32 on FP32 scratch.
template <typename S>
struct DafBuf {
    S* sd;    // dtau * spr of the sweep direction
    S* d[4];  // sweep right-hand sides, then d'
    S* c[4];  // c'
    S* x[4];  // sweep solution
};

// ---- pointwise transform before a sweep (rhs_diag_solver.F90:56-122,
// :300-311, :476-487) ----
// For every node of the level: sd = dtau * spr_D, the only per-node input
// the tridiagonal rows need.  For interior nodes: the right-hand side of the
// sweep, written to B.d:
//   D = 0: Ma^-1 S (-dtau J rh),  D = 1: N1^-1 rh,  D = 2: N2^-1 rh.
// The thread at the first (last) interior node of a line also builds the
// semi-explicit boundary vector brh of that line side (fine level only)
// from the untransformed rh, into F.brh[line * 8 + side * 4 + m].
template <int D, typename S>
__global__ void __launch_bounds__(256) k_daf_pre(Level L, Fields F, Consts C, DafBuf<S> B, int use_b1, int use_b2,
                                                 int b2_full, int ext) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= L.n) return;
    const int ni = L.ni, nj = L.nj, nk = L.nk;
    const int i = (int)(t % ni), j = (int)((t / ni) % nj), k = (int)(t / ((long)ni * nj));
    const long g = L.off + t;
    B.sd[t] = (S)(F.dtau[g] * spr_at<D>(F, C, g));
    // ext: the characteristic-outlet plane i = ni - 1 is solved as well
    if (i == 0 || i >= ni - 1 + ext || j == 0 || j == nj - 1 || k == 0 || k == nk - 1) return;
    const long sI = 1, sJ = ni, sK = (long)ni * nj;
    const double alpha_mult = C.e_source * 1.5 / C.delti;
    // S (-dtau J rh) at a node (:56-72)
    auto scaled = [&](long gg, double v[4]) {
        const double dt = F.dtau[gg];
        const double alpha = 1.0 + alpha_mult * dt;
        double r[4];
#pragma unroll
        for (int m = 0; m < 4; m++) r[m] = -dt * F.rh[m][gg] * F.aj[gg];
        v[0] = C.beta * r[0];
#pragma unroll
        for (int m = 1; m < 4; m++) v[m] = 1.0 / alpha * r[m];
    };
    // input of sweeps j and k: the previous sweep's solution
    auto xin = [&](int m, long tt) { return (double)B.x[m][tt]; };
    double v[4];
    if (D == 0) {
        scaled(g, v);
        apply_mai(F, C, g, v);
    } else {
#pragma unroll
        for (int m = 0; m < 4; m++) v[m] = xin(m, t);
        if (D == 1) apply_n1i(F, C, g, v);
        else apply_n2i(F, C, g, v);
    }
#pragma unroll
    for (int m = 0; m < 4; m++) B.d[m][t] = (S)v[m];
    if (!use_b1 && !use_b2) return;
    const int p = D == 0 ? i : D == 1 ? j : k;
    const int nd = D == 0 ? ni : D == 1 ? nj : nk;
    const long sD = D == 0 ? sI : D == 1 ? sJ : sK;
    const long line = D == 0 ? (long)(j - 1) + (long)(nj - 2) * (k - 1)
                    : D == 1 ? (long)(i - 1) + (long)(ni - 2 + ext) * (k - 1)
                             : (long)(i - 1) + (long)(ni - 2 + ext) * (j - 1);
    double* brh = F.brh + line * 8;
    const int b0 = 2 * D, b1 = 2 * D + 1;
    if (p == 1 && use_b1) {
        double r1[4], r2[4], w[4];
        if (D == 0) {
            double r3[4];
            scaled(g, r1);
            scaled(g + sD, r2);
            scaled(g + 2 * sD, r3);
            // fifth product of sa(1:5,1)*rh(1:4,..) (:97): lands in brh(1,2,j,k)
            const double ovf = C.sa[0][4] * r2[0] + C.sb[0][4] * r3[0];
            brh[4] = ovf; brh[5] = 0.0; brh[6] = 0.0; brh[7] = 0.0;
        } else {
#pragma unroll
            for (int m = 0; m < 4; m++) { r1[m] = xin(m, t); r2[m] = xin(m, t + sD); }
        }
#pragma unroll
        for (int m = 0; m < 4; m++) w[m] = C.sa[b0][m] * r1[m] + C.sb[b0][m] * r2[m];
        if (D == 0) apply_mai(F, C, g - sD, w);
        else if (D == 1) apply_n1i(F, C, g - sD, w);
        else apply_n2i(F, C, g - sD, w);
#pragma unroll
        for (int m = 0; m < 4; m++) brh[m] = w[m];
        if (D == 2 && !b2_full && use_b2) {
            // brh(:,2) is not rebuilt when btype(6) == 6 (:467) and keeps the
            // fifth product of the k = 1 build (:462), which reads rh(1) of node i+1
            double o[4] = {C.sa[4][4] * xin(0, t + sI) + C.sb[4][4] * xin(0, t + sI + sD), 0.0, 0.0, 0.0};
            apply_n2i(F, C, g + (long)(nd - 2) * sD, o);
#pragma unroll
            for (int m = 0; m < 4; m++) brh[4 + m] = o[m];
        }
    }
    if (p == nd - 2 && use_b2 && (D == 1 || (D == 2 && b2_full))) {
        double w[4];
#pragma unroll
        for (int m = 0; m < 4; m++) w[m] = C.sa[b1][m] * xin(m, t) + C.sb[b1][m] * xin(m, t - sD);
        if (D == 1) apply_n1i(F, C, g + sD, w);
        else apply_n2i(F, C, g + sD, w);
#pragma unroll
        for (int m = 0; m < 4; m++) brh[4 + m] = w[m];
    }
}

// ---- one scalar tridiagonal system per thread (rhs_diag_solver.F90:149-249
// and its j/k copies, the four SGTSV calls of a line) ----
// blockIdx.y = system m (eigenvalue multiplier 0, 0, +1, -1).  Rows are
// rebuilt from sd = dtau spr of the node and its two neighbours:
//   tmp = sd_s ep h^2, dds = dtc sd_(s-+1) h/2,
//   a = -lambda dds_w - tmp, b = 1 + 2 tmp, c = lambda dds_e - tmp.
// Thomas elimination in type S: c' in B.c[m], d' over the input in B.d[m],
// the solution into B.x[m].  Loads run one group of four rows ahead of the
// elimination.
// rows loaded per group, one group ahead of the elimination (16 measured
// best among 4, 8 and 16 at 129^3)
constexpr int kPF = 16;

template <int D, typename S>
__global__ void __launch_bounds__(128) k_daf_tri(Level L, Fields F, Consts C, DafBuf<S> B, int use_b1, int use_b2,
                                                 int ext) {
    const int ni = L.ni, nj = L.nj, nk = L.nk;
    const int nI = ni - 2 + ext;  // i positions of the j and k lines (ext: the outlet plane too)
    const int m = blockIdx.y;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long nlines, base, sD;
    int len;
    if (D == 0) {
        nlines = (long)(nj - 2) * (nk - 2);
        if (t >= nlines) return;
        base = (long)ni * ((1 + t % (nj - 2)) + (long)nj * (1 + t / (nj - 2)));
        sD = 1;
        len = ni - 2;
    } else if (D == 1) {
        nlines = (long)nI * (nk - 2);
        if (t >= nlines) return;
        base = (1 + t % nI) + (long)ni * nj * (1 + t / nI);
        sD = ni;
        len = nj - 2;
    } else {
        nlines = (long)nI * (nj - 2);
        if (t >= nlines) return;
        base = (1 + t % nI) + (long)ni * (1 + t / nI);
        sD = (long)ni * nj;
        len = nk - 2;
    }
    const S lam = m == 2 ? S(1) : m == 3 ? S(-1) : S(0);
    const S h = (S)(D == 0 ? L.dc : D == 1 ? L.de : L.dz);
    const S h2 = S(0.5) * h, hsq = h * h, epd = (S)C.ep[0][D];
    const S* __restrict__ sd = B.sd;
    S* __restrict__ dpm = B.d[m];
    S* __restrict__ cpm = B.c[m];
    S* __restrict__ rhm = B.x[m];
    const S brh1 = use_b1 ? (S)F.brh[t * 8 + m] : S(0);
    const S brh2 = use_b2 ? (S)F.brh[t * 8 + 4 + m] : S(0);

    S sdm = sd[base], sd0 = sd[base + sD];
    S cprev = 0, dprev = 0, denprev = 1;
    unsigned piv = 0;
    S dcur[kPF], scur[kPF];
#pragma unroll
    for (int u = 0; u < kPF; u++) {
        const long p = base + (1 + u) * sD;
        dcur[u] = u < len ? dpm[p] : S(0);
        scur[u] = u < len ? sd[p + sD] : S(0);
    }
    for (int s0 = 0; s0 < len; s0 += kPF) {
        S dnx[kPF], snx[kPF];
#pragma unroll
        for (int u = 0; u < kPF; u++) {
            const int s = s0 + kPF + u;
            const long p = base + (1 + s) * sD;
            dnx[u] = s < len ? dpm[p] : S(0);
            snx[u] = s < len ? sd[p + sD] : S(0);
        }
#pragma unroll
        for (int u = 0; u < kPF; u++) {
            const int s = s0 + u;
            if (s < len) {
                const long p = base + (1 + s) * sD;
                const S sdp = scur[u];
                const S tmp = sd0 * epd * hsq;
                const S dtc = (sd0 == S(0)) ? S(0) : S(1);
                const S ddw = dtc * sdm * h2, dde = dtc * sdp * h2;
                const S a = -ddw * lam - tmp;
                const S c = dde * lam - tmp;
                const S b = S(1) + S(2) * tmp;
                S d = dcur[u];
                if (s == 0 && use_b1) d = d - a * brh1;
                if (s == len - 1 && use_b2) d = d - c * brh2;
                S den;
                if (s == 0) {
                    den = b;
                } else {
                    if (fabs(denprev) < fabs(a)) piv++;  // SGTSV would interchange rows here
                    den = b - a * cprev;
                }
                const S inv = S(1) / den;
                cprev = c * inv;
                dprev = (s == 0) ? d * inv : (d - a * dprev) * inv;
                denprev = den;
                cpm[p] = cprev;
                dpm[p] = dprev;
                sdm = sd0;
                sd0 = sdp;
            }
        }
#pragma unroll
        for (int u = 0; u < kPF; u++) { dcur[u] = dnx[u]; scur[u] = snx[u]; }
    }
    // back substitution, loads one group ahead
    S x = dprev;
    rhm[base + (long)len * sD] = x;
    for (int s0 = len - 2; s0 >= 0; s0 -= kPF) {
        S dd[kPF], cc[kPF];
#pragma unroll
        for (int u = 0; u < kPF; u++) {
            const int s = s0 - u;
            const long p = base + (1 + s) * sD;
            dd[u] = s >= 0 ? dpm[p] : S(0);
            cc[u] = s >= 0 ? cpm[p] : S(0);
        }
#pragma unroll
        for (int u = 0; u < kPF; u++) {
            const int s = s0 - u;
            if (s >= 0) {
                x = dd[u] - cc[u] * x;
                rhm[base + (1 + s) * sD] = x;
            }
        }
    }
    if (piv) atomicAdd(F.pivots, (unsigned long long)piv);
}

// Two-way form of k_daf_tri for short lines: each wave solves 32 (line,
// system) pairs, lanes 0-31 eliminating rows 0 .. m-1 downwards with the
// Thomas recurrence above, lanes 32-63 rows len-1 .. m+1 upwards with the
// mirrored one (e' = a / den, f' = (d - c f') / den).  They meet at row m
// = len / 2 through a lane shuffle and substitute outwards.  On the coarse
// levels, where the one-way sweep leaves most CUs idle, this doubles the
// waves and halves the serial chain.  Rows below m are eliminated in the
// other order (a rounding-level difference); SGTSV pivot events are counted
// on the upper halves.  Needs len >= 4.
template <int D, typename S>
__global__ void __launch_bounds__(128) k_daf_tri2(Level L, Fields F, Consts C, DafBuf<S> B, int use_b1, int use_b2) {
    const int ni = L.ni, nj = L.nj, nk = L.nk;
    const int m = blockIdx.y;
    const int lane = (int)(threadIdx.x & 63), wave = (int)(threadIdx.x >> 6);
    const bool top = lane < 32;
    const long nlines = D == 0 ? (long)(nj - 2) * (nk - 2) : D == 1 ? (long)(ni - 2) * (nk - 2)
                                                                    : (long)(ni - 2) * (nj - 2);
    const long t = (long)blockIdx.x * 64 + wave * 32 + (lane & 31);
    const bool act = t < nlines;
    const long sD = D == 0 ? 1 : D == 1 ? (long)ni : (long)ni * nj;
    const int len = D == 0 ? ni - 2 : D == 1 ? nj - 2 : nk - 2;
    long base = 0;
    if (act) {
        if (D == 0) base = (long)ni * ((1 + t % (nj - 2)) + (long)nj * (1 + t / (nj - 2)));
        else if (D == 1) base = (1 + t % (ni - 2)) + (long)ni * nj * (1 + t / (ni - 2));
        else base = (1 + t % (ni - 2)) + (long)ni * (1 + t / (ni - 2));
    }
    const S lam = m == 2 ? S(1) : m == 3 ? S(-1) : S(0);
    const S h = (S)(D == 0 ? L.dc : D == 1 ? L.de : L.dz);
    const S h2 = S(0.5) * h, hsq = h * h, epd = (S)C.ep[0][D];
    const S* __restrict__ sd = B.sd;
    S* __restrict__ dpm = B.d[m];
    S* __restrict__ cpm = B.c[m];
    S* __restrict__ rhm = B.x[m];
    const S brh1 = (use_b1 && act) ? (S)F.brh[t * 8 + m] : S(0);
    const S brh2 = (use_b2 && act) ? (S)F.brh[t * 8 + 4 + m] : S(0);
    const int mid = len / 2;
    auto node = [&](int pos) { return base + (long)pos * sD; };
    // upper half carries sd(s), sd(s+1); lower half sd(s+1), sd(s+2)
    S s_a = 0, s_b = 0;
    if (act) {
        if (top) { s_a = sd[node(0)]; s_b = sd[node(1)]; }
        else { s_a = sd[node(len)]; s_b = sd[node(len + 1)]; }
    }
    auto coef = [&](S sdm, S sd0, S sdp, S& a, S& b, S& c) {
        const S tmp = sd0 * epd * hsq;
        const S dtc = (sd0 == S(0)) ? S(0) : S(1);
        const S ddw = dtc * sdm * h2, dde = dtc * sdp * h2;
        a = -ddw * lam - tmp;
        c = dde * lam - tmp;
        b = S(1) + S(2) * tmp;
    };
    S cprev = 0, dprev = 0, denprev = 1;
    unsigned piv = 0;
    const int nrows = top ? mid : len - 1 - mid;
    for (int g0 = 0; g0 < nrows; g0 += kPF) {
        S ls[kPF], ld[kPF];
#pragma unroll
        for (int u = 0; u < kPF; u++) {
            const int r = g0 + u;
            const int srow = top ? r : len - 1 - r;
            const bool ok = act && r < nrows;
            ls[u] = ok ? sd[top ? node(srow + 2) : node(srow)] : S(0);
            ld[u] = ok ? dpm[node(srow + 1)] : S(0);
        }
#pragma unroll
        for (int u = 0; u < kPF; u++) {
            const int r = g0 + u;
            if (r < nrows) {
                const int srow = top ? r : len - 1 - r;
                S sdm, sd0, sdp;
                if (top) { sdm = s_a; sd0 = s_b; sdp = ls[u]; }
                else { sd0 = s_a; sdp = s_b; sdm = ls[u]; }
                S a, b, c;
                coef(sdm, sd0, sdp, a, b, c);
                S d = ld[u];
                if (srow == 0 && use_b1) d = d - a * brh1;
                if (srow == len - 1 && use_b2) d = d - c * brh2;
                S den;
                if (r == 0) {
                    den = b;
                } else {
                    if (top && fabs(denprev) < fabs(a)) piv++;  // SGTSV would interchange rows here
                    den = top ? b - a * cprev : b - c * cprev;
                }
                const S inv = S(1) / den;
                if (top) {
                    cprev = c * inv;
                    dprev = (r == 0) ? d * inv : (d - a * dprev) * inv;
                } else {
                    cprev = a * inv;
                    dprev = (r == 0) ? d * inv : (d - c * dprev) * inv;
                }
                denprev = den;
                if (act) {
                    cpm[node(srow + 1)] = cprev;
                    dpm[node(srow + 1)] = dprev;
                }
                if (top) { s_a = sd0; s_b = sdp; }
                else { s_b = sd0; s_a = sdm; }
            }
        }
    }
    const S e_lo = __shfl(cprev, (lane & 31) + 32, 64);
    const S f_lo = __shfl(dprev, (lane & 31) + 32, 64);
    S xm = 0;
    if (top && act) {
        S a, b, c;
        coef(s_a, s_b, sd[node(mid + 2)], a, b, c);
        const S d = dpm[node(mid + 1)];
        if (fabs(denprev) < fabs(a)) piv++;
        xm = (d - a * dprev - c * f_lo) / (b - a * cprev - c * e_lo);
        rhm[node(mid + 1)] = xm;
    }
    xm = __shfl(xm, lane & 31, 64);
    S x = xm;
    const int nb = top ? mid : len - 1 - mid;
    for (int g0 = 0; g0 < nb; g0 += kPF) {
        S dd[kPF], cc[kPF];
#pragma unroll
        for (int u = 0; u < kPF; u++) {
            const int r = g0 + u;
            const int srow = top ? mid - 1 - r : mid + 1 + r;
            const bool ok = act && r < nb;
            dd[u] = ok ? dpm[node(srow + 1)] : S(0);
            cc[u] = ok ? cpm[node(srow + 1)] : S(0);
        }
#pragma unroll
        for (int u = 0; u < kPF; u++) {
            const int r = g0 + u;
            if (r < nb && act) {
                const int srow = top ? mid - 1 - r : mid + 1 + r;
                x = dd[u] - cc[u] * x;
                rhm[node(srow + 1)] = x;
            }
        }
    }
    if (piv) atomicAdd(F.pivots, (unsigned long long)piv);
}

// The i-direction sweep of k_daf_tri with its rows staged through LDS
// (line_tile.h): a workgroup owns kTL i-lines, thread = (system m, line l).
// Chunks of kTR rows of sd and d are loaded one chunk ahead of the
// elimination; c' and d' go back to HBM per chunk except for the last chunk,
// which the back substitution starts from in LDS.  Same arithmetic as
// k_daf_tri<0>.
template <typename S>
__global__ void __launch_bounds__(4 * kTL) k_daf_tri_i(Level L, Fields F, Consts C, DafBuf<S> B, int use_b1,
                                                       int use_b2, int ext) {
    constexpr int NT = 4 * kTL;
    constexpr int kSS = kTR + 3;  // LDS stride of the sd tile (kTR + 2 rows, odd)
    using IOr = TileIO<NT, kTL, kTR, kTS, S>;
    using IOs = TileIO<NT, kTL, kTR + 2, kSS, S>;
    __shared__ long s_base[kTL];
    __shared__ S s_sd[kTL * kSS];
    __shared__ S s_d[4][kTL * kTS];
    __shared__ S s_c[4][kTL * kTS];

    const long nlines = (long)(L.nj - 2) * (L.nk - 2);
    const long t0 = (long)blockIdx.x * kTL;
    const int nl = (int)min((long)kTL, nlines - t0);
    const int m = (int)threadIdx.x / kTL, l = (int)threadIdx.x % kTL;
    const bool act = l < nl;
    const long t = t0 + l;
    if (threadIdx.x < kTL) s_base[threadIdx.x] = act ? iline_base(L, t) : 0;
    __syncthreads();

    // ext: the line closes on the characteristic-outlet plane i = ni - 1
    const int len = L.ni - 2 + ext;
    const int nsd = L.ni;  // sd rows loaded: nodes i = 0 .. ni - 1 of the line
    const int nch = (len + kTR - 1) / kTR;
    const S lam = m == 2 ? S(1) : m == 3 ? S(-1) : S(0);
    const S h = (S)L.dc;
    const S h2 = S(0.5) * h, hsq = h * h, epd = (S)C.ep[0][0];
    const S* __restrict__ sd = B.sd;
    const S brh1 = (use_b1 && act) ? (S)F.brh[t * 8 + m] : S(0);
    const S brh2 = (use_b2 && act) ? (S)F.brh[t * 8 + 4 + m] : S(0);
    S* myd = s_d[m] + l * kTS;
    S* myc = s_c[m] + l * kTS;
    const S* mys = s_sd + l * kSS;

    // row s of a line is node base + 1 + s; sd of rows s-1..s+1 are nodes base + s .. base + s + 2
    S vd[4][IOr::kPer], vs[IOs::kPer];
    for (int mm = 0; mm < 4; mm++) IOr::load(vd[mm], B.d[mm] + 1, s_base, nl, 0, len);
    IOs::load(vs, sd, s_base, nl, 0, nsd);

    S cprev = 0, dprev = 0, denprev = 1;
    unsigned piv = 0;
    for (int ch = 0; ch < nch; ch++) {
        const int s0 = ch * kTR;
        __syncthreads();
        for (int mm = 0; mm < 4; mm++) IOr::to_lds(vd[mm], s_d[mm]);
        IOs::to_lds(vs, s_sd);
        __syncthreads();
        if (ch + 1 < nch) {
            for (int mm = 0; mm < 4; mm++) IOr::load(vd[mm], B.d[mm] + 1, s_base, nl, s0 + kTR, len);
            IOs::load(vs, sd, s_base, nl, s0 + kTR, nsd);
        }
        if (act) {
            for (int u = 0; u < kTR; u++) {
                const int s = s0 + u;
                if (s >= len) break;
                const S sdm = mys[u], sd0 = mys[u + 1], sdp = mys[u + 2];
                const S tmp = sd0 * epd * hsq;
                const S dtc = (sd0 == S(0)) ? S(0) : S(1);
                const S ddw = dtc * sdm * h2, dde = dtc * sdp * h2;
                S a = -ddw * lam - tmp;
                S c = dde * lam - tmp;
                S b = S(1) + S(2) * tmp;
                if (ext && s == len - 1) {
                    // outlet row (rhs_diag_solver.F90:207-221): upwind coupling
                    // of the +1 eigenvalue only, no smoothing
                    a = m == 2 ? -(sdm * h) : S(0);
                    b = m == 2 ? S(1) + sd0 * h : S(1);
                    c = S(0);
                }
                S d = myd[u];
                if (s == 0 && use_b1) d = d - a * brh1;
                if (s == len - 1 && use_b2) d = d - c * brh2;
                S den;
                if (s == 0) {
                    den = b;
                } else {
                    if (fabs(denprev) < fabs(a)) piv++;  // SGTSV would interchange rows here
                    den = b - a * cprev;
                }
                const S inv = S(1) / den;
                cprev = c * inv;
                dprev = (s == 0) ? d * inv : (d - a * dprev) * inv;
                denprev = den;
                myc[u] = cprev;
                myd[u] = dprev;
            }
        }
        __syncthreads();
        if (ch + 1 < nch) {
            for (int mm = 0; mm < 4; mm++) {
                IOr::store(s_c[mm], B.c[mm] + 1, s_base, nl, s0, len);
                IOr::store(s_d[mm], B.d[mm] + 1, s_base, nl, s0, len);
            }
        }
    }
    // back substitution, last chunk first (already in LDS)
    S vc[4][IOr::kPer];
    S x = 0;
    for (int ch = nch - 1; ch >= 0; ch--) {
        const int s0 = ch * kTR;
        if (ch < nch - 1) {
            __syncthreads();
            for (int mm = 0; mm < 4; mm++) {
                IOr::to_lds(vc[mm], s_c[mm]);
                IOr::to_lds(vd[mm], s_d[mm]);
            }
            __syncthreads();
        }
        if (ch > 0) {
            for (int mm = 0; mm < 4; mm++) {
                IOr::load(vc[mm], B.c[mm] + 1, s_base, nl, s0 - kTR, len);
                IOr::load(vd[mm], B.d[mm] + 1, s_base, nl, s0 - kTR, len);
            }
        }
        if (act) {
            for (int u = kTR - 1; u >= 0; u--) {
                const int s = s0 + u;
                if (s >= len) continue;
                x = (s == len - 1) ? myd[u] : myd[u] - myc[u] * x;
                myd[u] = x;
            }
        }
        __syncthreads();
        for (int mm = 0; mm < 4; mm++) IOr::store(s_d[mm], B.x[mm] + 1, s_base, nl, s0, len);
    }
    if (piv) atomicAdd(F.pivots, (unsigned long long)piv);
}

// dQ = Mc dQ*** and q = q + dQ on interior nodes (rhs_diag_solver.F90:624-633,
// solver_daf.F90:409-418)
template <typename S>
__global__ void __launch_bounds__(256) k_daf_post(Level L, Fields F, Consts C, DafBuf<S> B, double* dq_out, int ext) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= L.n) return;
    const int ni = L.ni, nj = L.nj, nk = L.nk;
    const int i = (int)(t % ni), j = (int)((t / ni) % nj), k = (int)(t / ((long)ni * nj));
    if (i == 0 || i >= ni - 1 + ext || j == 0 || j == nj - 1 || k == 0 || k == nk - 1) return;
    const long g = L.off + t;
    double v[4];
#pragma unroll
    for (int m = 0; m < 4; m++) v[m] = (double)B.x[m][t];
    apply_mc(F, C, g, v);
#pragma unroll
    for (int m = 0; m < 4; m++) F.q[m][g] = F.q[m][g] + v[m];
    if (dq_out) {
#pragma unroll
        for (int m = 0; m < 4; m++) dq_out[m * L.n + t] = v[m];
    }
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

template <typename S>
static void daf_sweep(hipStream_t s, const Level& L, const Fields& F, const Consts& C, const DafBuf<S>& B, int n,
                      int D) {
    const int use_b1 = (n == 1 && C.btype[2 * D] != 5);
    const int use_b2 = (n == 1 && C.btype[2 * D + 1] != 5);
    const int b2_full = (n == 1 && C.btype[5] != 6);
    // a characteristic outlet at i = ni - 1 joins the fine level's sweeps
    const int ext = (n == 1 && C.btype[1] == DESALM_BC_OUTLET_MOC);
    const unsigned gp = (unsigned)((L.n + 255) / 256);
    const long nl = D == 0 ? (long)(L.nj - 2) * (L.nk - 2) : D == 1 ? (long)(L.ni - 2 + ext) * (L.nk - 2)
                                                                    : (long)(L.ni - 2 + ext) * (L.nj - 2);
    const dim3 gt((unsigned)((nl + 127) / 128), 4);
    const int len = D == 0 ? L.ni - 2 : D == 1 ? L.nj - 2 : L.nk - 2;
    if (n > 1 && len >= 4) {
        // coarse levels: the two-way sweep
        const dim3 g2((unsigned)((nl + 63) / 64), 4);
        if (D == 0) {
            hipLaunchKernelGGL((k_daf_pre<0, S>), dim3(gp), dim3(256), 0, s, L, F, C, B, use_b1, use_b2, b2_full, ext);
            hipLaunchKernelGGL((k_daf_tri2<0, S>), g2, dim3(128), 0, s, L, F, C, B, use_b1, use_b2);
        } else if (D == 1) {
            hipLaunchKernelGGL((k_daf_pre<1, S>), dim3(gp), dim3(256), 0, s, L, F, C, B, use_b1, use_b2, b2_full, ext);
            hipLaunchKernelGGL((k_daf_tri2<1, S>), g2, dim3(128), 0, s, L, F, C, B, use_b1, use_b2);
        } else {
            hipLaunchKernelGGL((k_daf_pre<2, S>), dim3(gp), dim3(256), 0, s, L, F, C, B, use_b1, use_b2, b2_full, ext);
            hipLaunchKernelGGL((k_daf_tri2<2, S>), g2, dim3(128), 0, s, L, F, C, B, use_b1, use_b2);
        }
        return;
    }
    if (D == 0) {
        hipLaunchKernelGGL((k_daf_pre<0, S>), dim3(gp), dim3(256), 0, s, L, F, C, B, use_b1, use_b2, b2_full, ext);
        const unsigned gi = (unsigned)((nl + kTL - 1) / kTL);
        hipLaunchKernelGGL(k_daf_tri_i<S>, dim3(gi), dim3(4 * kTL), 0, s, L, F, C, B, use_b1, use_b2, ext);
    } else if (D == 1) {
        hipLaunchKernelGGL((k_daf_pre<1, S>), dim3(gp), dim3(256), 0, s, L, F, C, B, use_b1, use_b2, b2_full, ext);
        hipLaunchKernelGGL((k_daf_tri<1, S>), gt, dim3(128), 0, s, L, F, C, B, use_b1, use_b2, ext);
    } else {
        hipLaunchKernelGGL((k_daf_pre<2, S>), dim3(gp), dim3(256), 0, s, L, F, C, B, use_b1, use_b2, b2_full, ext);
        hipLaunchKernelGGL((k_daf_tri<2, S>), gt, dim3(128), 0, s, L, F, C, B, use_b1, use_b2, ext);
    }
}

static DafBuf<double> buf64(const Level& L, const Fields& F) {
    DafBuf<double> B;
    B.sd = F.sd;
    for (int m = 0; m < 4; m++) {
        B.d[m] = F.dp[m];
        B.c[m] = F.cp[m];
        B.x[m] = F.rh[m] + L.off;
    }
    return B;
}

static DafBuf<float> buf32(const Fields& F) {
    DafBuf<float> B;
    B.sd = F.sdf;
    for (int m = 0; m < 4; m++) {
        B.d[m] = F.dpf[m];
        B.c[m] = F.cpf[m];
        B.x[m] = F.xf[m];
    }
    return B;
}

static void daf_dir(hipStream_t s, const Level& L, const Fields& F, const Consts& C, int n, int D) {
    if (C.mixed) daf_sweep(s, L, F, C, buf32(F), n, D);
    else daf_sweep(s, L, F, C, buf64(L, F), n, D);
}

void launch_daf_i(hipStream_t s, const Level& L, const Fields& F, const Consts& C, int n) { daf_dir(s, L, F, C, n, 0); }

void launch_daf_j(hipStream_t s, const Level& L, const Fields& F, const Consts& C, int n) { daf_dir(s, L, F, C, n, 1); }

void launch_daf_k(hipStream_t s, const Level& L, const Fields& F, const Consts& C, int n, double* dq_out) {
    daf_dir(s, L, F, C, n, 2);
    const dim3 g((unsigned)((L.n + 255) / 256));
    const int ext = (n == 1 && C.btype[1] == DESALM_BC_OUTLET_MOC);
    if (C.mixed) hipLaunchKernelGGL(k_daf_post<float>, g, dim3(256), 0, s, L, F, C, buf32(F), dq_out, ext);
    else hipLaunchKernelGGL(k_daf_post<double>, g, dim3(256), 0, s, L, F, C, buf64(L, F), dq_out, ext);
}

void launch_gtsv_batch(hipStream_t s, int n, int nsys, double* dl, double* d, double* du, double* b, int* info) {
    const int bs = 64;
    hipLaunchKernelGGL(k_gtsv_batch, dim3((unsigned)((nsys + bs - 1) / bs)), dim3(bs), 0, s, n, nsys, dl, d, du, b,
                       info);
}

}  // namespace desalm
