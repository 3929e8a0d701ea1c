// k_daf.hip -- diagonalized approximate-factorization implicit operator.
//
// rhs_modal_matrices + rhs_diag_solver (src/rhs_modal_matrices.F90,
// src/rhs_diag_solver.F90) as three line-sweep kernels:
//   i sweep: S scaling, Ma^-1, i-direction tridiagonal solves
//   j sweep: N1^-1 = Mb^-1 Ma, j-direction solves
//   k sweep: N2^-1 = Mc^-1 Mb, k-direction solves, Mc, and q = q + dq
// The eigenvector matrices are rebuilt in registers from the metrics and
// dtau of each node instead of being stored (the reference keeps 67 words
// per node).  One thread owns one grid line and runs a Thomas elimination
// on its four scalar systems (three distinct matrices: eigenvalue
// multipliers 0, 0, +1, -1).  LAPACK SGTSV, which the reference calls, would
// pick the same pivots as Thomas unless |d_i| < |dl_i|; rows where that
// happens are counted in F.pivots so a caller can see when the pivot-free
// elimination departed from SGTSV's row order.
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
template <int D>
__device__ __forceinline__ double spr_at(const Fields& F, const Consts& C, long g) {
    const double* m0 = D == 0 ? F.csi[0] : D == 1 ? F.eta[0] : F.zet[0];
    const double* m1 = D == 0 ? F.csi[1] : D == 1 ? F.eta[1] : F.zet[1];
    const double* m2 = D == 0 ? F.csi[2] : D == 1 ? F.eta[2] : F.zet[2];
    const double gg = m0[g] * m0[g] + m1[g] * m1[g] + m2[g] * m2[g];
    return modal_theta(C, F.dtau[g]) * sqrt(gg);
}

// Line-sweep state for the three distinct DAF matrices
struct Thomas3 {
    double cp[3];   // c' of matrices 0 (lambda 0), 1 (lambda +1), 2 (lambda -1)
    double dp[4];   // d' of the four systems
    double den[3];  // modified diagonal of the previous row
};

// one forward row: coefficients (a_X, b, c_X), right-hand sides aq
__device__ __forceinline__ void thomas_row(Thomas3& T, bool first, const double a[3], double b, const double c[3],
                                           const double aq[4], unsigned& piv) {
    double inv[3];
#pragma unroll
    for (int x = 0; x < 3; x++) {
        double den;
        if (first) {
            den = b;
        } else {
            // SGTSV would interchange rows when |d_prev| < |dl|
            if (fabs(T.den[x]) < fabs(a[x])) piv++;
            den = b - a[x] * T.cp[x];
        }
        T.den[x] = den;
        inv[x] = 1.0 / den;
        T.cp[x] = c[x] * inv[x];
    }
    const int mat[4] = {0, 0, 1, 2};
#pragma unroll
    for (int m = 0; m < 4; m++) {
        const int x = mat[m];
        T.dp[m] = first ? aq[m] * inv[x] : (aq[m] - a[x] * T.dp[m]) * inv[x];
    }
}

// coefficients of one DAF row in direction D (rhs_diag_solver.F90:152-187)
__device__ __forceinline__ void daf_coefs(double dtau0, double spr0, double dtaum, double sprm, double dtaup,
                                          double sprp, double epd, double h, double a[3], double& b,
                                          double c[3]) {
    const double h2 = 0.5 * h, hsq = h * h;
    const double dtc = (dtau0 == 0.0) ? 0.0 : 1.0;
    const double tmp = dtau0 * epd * spr0 * hsq;
    b = 1.0 + 2.0 * tmp;
    const double ddw = dtc * dtaum * h2 * sprm;
    const double dde = dtc * dtaup * h2 * sprp;
    a[0] = -tmp;
    a[1] = -ddw * 1.0 - tmp;
    a[2] = -ddw * -1.0 - tmp;
    c[0] = -tmp;
    c[1] = dde * 1.0 - tmp;
    c[2] = dde * -1.0 - tmp;
}

// ---- i sweep (rhs_diag_solver.F90:51-258) ----
__global__ void __launch_bounds__(128) k_daf_i(Level L, Fields F, Consts C, int use_b1, int use_b2) {
    const int ni = L.ni, nj = L.nj, nk = L.nk;
    const long nlines = (long)(nj - 2) * (nk - 2);
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nlines) return;
    const int j = 1 + (int)(t % (nj - 2));
    const int k = 1 + (int)(t / (nj - 2));
    const long base = L.off + (long)ni * (j + (long)nj * k);  // global index of (0, j, k)
    const long lbase = (long)ni * (j + (long)nj * k);          // level-local index of (0, j, k)
    const int len = ni - 2;
    const double alpha_mult = C.e_source * 1.5 / C.delti;
    const double epd = C.ep[0][0];
    const double h = L.dc;

    // S (-dtau J rh) at a node (:56-72)
    auto scaled = [&](long g, double v[4]) {
        const double dt = F.dtau[g];
        const double alpha = 1.0 + alpha_mult * dt;
        double r[4];
#pragma unroll
        for (int m = 0; m < 4; m++) r[m] = -dt * F.rh[m][g] * F.aj[g];
        v[0] = C.beta * r[0];
#pragma unroll
        for (int m = 1; m < 4; m++) v[m] = 1.0 / alpha * r[m];
    };

    double brh1[4] = {0, 0, 0, 0};
    double ovf = 0.0;
    double s1[4], s2[4], s3[4];
    scaled(base + 1, s1);
    scaled(base + 2, s2);
    if (use_b1) {
        scaled(base + 3, s3);
#pragma unroll
        for (int m = 0; m < 4; m++) brh1[m] = C.sa[0][m] * s1[m] + C.sb[0][m] * s2[m];
        // fifth product of sa(1:5,1)*rh(1:4,..) (:97) lands in brh(1,2,j,k)
        ovf = C.sa[0][4] * s2[0] + C.sb[0][4] * s3[0];
        apply_mai(F, C, base, brh1);
    }

    Thomas3 T;
    unsigned piv = 0;
    double dtm = F.dtau[base], sprm = spr_at<0>(F, C, base);
    double dt0 = F.dtau[base + 1], spr0 = spr_at<0>(F, C, base + 1);
    for (int s = 0; s < len; s++) {
        const long g = base + 1 + s;
        const double dtp = F.dtau[g + 1], sprp = spr_at<0>(F, C, g + 1);
        double aq[4];
        if (s == 0) {
#pragma unroll
            for (int m = 0; m < 4; m++) aq[m] = s1[m];
        } else if (s == 1) {
#pragma unroll
            for (int m = 0; m < 4; m++) aq[m] = s2[m];
        } else {
            scaled(g, aq);
        }
        apply_mai(F, C, g, aq);
        double a[3], b, c[3];
        daf_coefs(dt0, spr0, dtm, sprm, dtp, sprp, epd, h, a, b, c);
        if (s == 0 && use_b1) {
            const double aw4[4] = {a[0], a[0], a[1], a[2]};
#pragma unroll
            for (int m = 0; m < 4; m++) aq[m] = aq[m] - aw4[m] * brh1[m];
        }
        if (s == len - 1 && use_b2) {
            // the i = imax correction is built only on interior ranks (:102)
            // and stays zero except for the overflowed component
            aq[0] = aq[0] - c[0] * ovf;
        }
        thomas_row(T, s == 0, a, b, c, aq, piv);
        const long lt = lbase + 1 + s;
#pragma unroll
        for (int x = 0; x < 3; x++) F.cp[x][lt] = T.cp[x];
#pragma unroll
        for (int m = 0; m < 4; m++) F.rh[m][g] = T.dp[m];
        dtm = dt0; sprm = spr0; dt0 = dtp; spr0 = sprp;
    }
    // back substitution
    double x[4];
#pragma unroll
    for (int m = 0; m < 4; m++) x[m] = T.dp[m];
    const int mat[4] = {0, 0, 1, 2};
    for (int s = len - 2; s >= 0; s--) {
        const long g = base + 1 + s;
        const long lt = lbase + 1 + s;
#pragma unroll
        for (int m = 0; m < 4; m++) {
            x[m] = F.rh[m][g] - F.cp[mat[m]][lt] * x[m];
            F.rh[m][g] = x[m];
        }
    }
    if (piv) atomicAdd(F.pivots, (unsigned long long)piv);
}

// ---- j sweep (rhs_diag_solver.F90:265-434) ----
__global__ void __launch_bounds__(128) k_daf_j(Level L, Fields F, Consts C, int use_b1, int use_b2) {
    const int ni = L.ni, nj = L.nj, nk = L.nk;
    const long nlines = (long)(ni - 2) * (nk - 2);
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nlines) return;
    const int i = 1 + (int)(t % (ni - 2));
    const int k = 1 + (int)(t / (ni - 2));
    const long sJ = ni;
    const long base = L.off + i + (long)ni * nj * k;  // (i, 0, k)
    const long lbase = i + (long)ni * nj * k;
    const int len = nj - 2;
    const double epd = C.ep[0][1];
    const double h = L.de;

    double brh1[4] = {0, 0, 0, 0}, brh2[4] = {0, 0, 0, 0};
    if (use_b1) {
#pragma unroll
        for (int m = 0; m < 4; m++) brh1[m] = C.sa[2][m] * F.rh[m][base + sJ] + C.sb[2][m] * F.rh[m][base + 2 * sJ];
        apply_n1i(F, C, base, brh1);
    }
    if (use_b2) {
        const long ge = base + (long)(nj - 2) * sJ;
#pragma unroll
        for (int m = 0; m < 4; m++) brh2[m] = C.sa[3][m] * F.rh[m][ge] + C.sb[3][m] * F.rh[m][ge - sJ];
        apply_n1i(F, C, base + (long)(nj - 1) * sJ, brh2);
    }

    Thomas3 T;
    unsigned piv = 0;
    double dtm = F.dtau[base], sprm = spr_at<1>(F, C, base);
    double dt0 = F.dtau[base + sJ], spr0 = spr_at<1>(F, C, base + sJ);
    for (int s = 0; s < len; s++) {
        const long g = base + (1 + s) * sJ;
        const double dtp = F.dtau[g + sJ], sprp = spr_at<1>(F, C, g + sJ);
        double aq[4];
#pragma unroll
        for (int m = 0; m < 4; m++) aq[m] = F.rh[m][g];
        apply_n1i(F, C, g, aq);
        double a[3], b, c[3];
        daf_coefs(dt0, spr0, dtm, sprm, dtp, sprp, epd, h, a, b, c);
        if (s == 0 && use_b1) {
            const double aw4[4] = {a[0], a[0], a[1], a[2]};
#pragma unroll
            for (int m = 0; m < 4; m++) aq[m] = aq[m] - aw4[m] * brh1[m];
        }
        if (s == len - 1 && use_b2) {
            const double ae4[4] = {c[0], c[0], c[1], c[2]};
#pragma unroll
            for (int m = 0; m < 4; m++) aq[m] = aq[m] - ae4[m] * brh2[m];
        }
        thomas_row(T, s == 0, a, b, c, aq, piv);
        const long lt = lbase + (1 + s) * sJ;
#pragma unroll
        for (int x = 0; x < 3; x++) F.cp[x][lt] = T.cp[x];
#pragma unroll
        for (int m = 0; m < 4; m++) F.rh[m][g] = T.dp[m];
        dtm = dt0; sprm = spr0; dt0 = dtp; spr0 = sprp;
    }
    double x[4];
#pragma unroll
    for (int m = 0; m < 4; m++) x[m] = T.dp[m];
    const int mat[4] = {0, 0, 1, 2};
    for (int s = len - 2; s >= 0; s--) {
        const long g = base + (1 + s) * sJ;
        const long lt = lbase + (1 + s) * sJ;
#pragma unroll
        for (int m = 0; m < 4; m++) {
            x[m] = F.rh[m][g] - F.cp[mat[m]][lt] * x[m];
            F.rh[m][g] = x[m];
        }
    }
    if (piv) atomicAdd(F.pivots, (unsigned long long)piv);
}

// ---- k sweep, Mc and the q update (rhs_diag_solver.F90:441-633,
// solver_daf.F90:409-418) ----
// rh is only read here (the k = 1 correction of a line also reads rh of
// the next i line), so d' goes to F.dp and the result straight into q.
__global__ void __launch_bounds__(128) k_daf_k(Level L, Fields F, Consts C, int use_b1, int build_b2_full,
                                               int use_b2, double* dq_out) {
    const int ni = L.ni, nj = L.nj, nk = L.nk;
    const long nlines = (long)(ni - 2) * (nj - 2);
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nlines) return;
    const int i = 1 + (int)(t % (ni - 2));
    const int j = 1 + (int)(t / (ni - 2));
    const long sK = (long)ni * nj;
    const long base = L.off + i + (long)ni * j;  // (i, j, 0)
    const long lbase = i + (long)ni * j;
    const int len = nk - 2;
    const double epd = C.ep[0][2];
    const double h = L.dz;

    double brh1[4] = {0, 0, 0, 0}, brh2[4] = {0, 0, 0, 0};
    if (use_b1) {
#pragma unroll
        for (int m = 0; m < 4; m++) brh1[m] = C.sa[4][m] * F.rh[m][base + sK] + C.sb[4][m] * F.rh[m][base + 2 * sK];
        // fifth product of sa(1:5,5)*rh(1:4,..) (:462) lands in brh(1,2,i,j);
        // it reads rh(1) of node i+1
        brh2[0] = C.sa[4][4] * F.rh[0][base + 1 + sK] + C.sb[4][4] * F.rh[0][base + 1 + 2 * sK];
        apply_n2i(F, C, base, brh1);
    }
    if (build_b2_full) {  // built under btype(6) /= 6 (:467)
        const long ge = base + (long)(nk - 2) * sK;
#pragma unroll
        for (int m = 0; m < 4; m++) brh2[m] = C.sa[5][m] * F.rh[m][ge] + C.sb[5][m] * F.rh[m][ge - sK];
    }
    if (use_b2) apply_n2i(F, C, base + (long)(nk - 1) * sK, brh2);  // transformed under /= 5 (:504)

    Thomas3 T;
    unsigned piv = 0;
    double dtm = F.dtau[base], sprm = spr_at<2>(F, C, base);
    double dt0 = F.dtau[base + sK], spr0 = spr_at<2>(F, C, base + sK);
    for (int s = 0; s < len; s++) {
        const long g = base + (1 + s) * sK;
        const double dtp = F.dtau[g + sK], sprp = spr_at<2>(F, C, g + sK);
        double aq[4];
#pragma unroll
        for (int m = 0; m < 4; m++) aq[m] = F.rh[m][g];
        apply_n2i(F, C, g, aq);
        double a[3], b, c[3];
        daf_coefs(dt0, spr0, dtm, sprm, dtp, sprp, epd, h, a, b, c);
        if (s == 0 && use_b1) {
            const double aw4[4] = {a[0], a[0], a[1], a[2]};
#pragma unroll
            for (int m = 0; m < 4; m++) aq[m] = aq[m] - aw4[m] * brh1[m];
        }
        if (s == len - 1 && use_b2) {
            const double ae4[4] = {c[0], c[0], c[1], c[2]};
#pragma unroll
            for (int m = 0; m < 4; m++) aq[m] = aq[m] - ae4[m] * brh2[m];
        }
        thomas_row(T, s == 0, a, b, c, aq, piv);
        const long lt = lbase + (1 + s) * sK;
#pragma unroll
        for (int x = 0; x < 3; x++) F.cp[x][lt] = T.cp[x];
#pragma unroll
        for (int m = 0; m < 4; m++) F.dp[m][lt] = T.dp[m];
        dtm = dt0; sprm = spr0; dt0 = dtp; spr0 = sprp;
    }
    const long nl = L.n;
    double x[4];
#pragma unroll
    for (int m = 0; m < 4; m++) x[m] = T.dp[m];
    const int mat[4] = {0, 0, 1, 2};
    for (int s = len - 1; s >= 0; s--) {
        const long g = base + (1 + s) * sK;
        const long lt = lbase + (1 + s) * sK;
        if (s < len - 1) {
#pragma unroll
            for (int m = 0; m < 4; m++) x[m] = F.dp[m][lt] - F.cp[mat[m]][lt] * x[m];
        }
        double dq[4] = {x[0], x[1], x[2], x[3]};
        apply_mc(F, C, g, dq);
#pragma unroll
        for (int m = 0; m < 4; m++) F.q[m][g] = F.q[m][g] + dq[m];
        if (dq_out) {
#pragma unroll
            for (int m = 0; m < 4; m++) dq_out[m * nl + lt] = dq[m];
        }
    }
    if (piv) atomicAdd(F.pivots, (unsigned long long)piv);
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

void launch_daf_i(hipStream_t s, const Level& L, const Fields& F, const Consts& C, int n) {
    const long nl = (long)(L.nj - 2) * (L.nk - 2);
    const int b = 128;
    const int use_b1 = (n == 1 && C.btype[0] != 5);
    const int use_b2 = (n == 1 && C.btype[1] != 5);
    hipLaunchKernelGGL(k_daf_i, dim3((unsigned)((nl + b - 1) / b)), dim3(b), 0, s, L, F, C, use_b1, use_b2);
}

void launch_daf_j(hipStream_t s, const Level& L, const Fields& F, const Consts& C, int n) {
    const long nl = (long)(L.ni - 2) * (L.nk - 2);
    const int b = 128;
    const int use_b1 = (n == 1 && C.btype[2] != 5);
    const int use_b2 = (n == 1 && C.btype[3] != 5);
    hipLaunchKernelGGL(k_daf_j, dim3((unsigned)((nl + b - 1) / b)), dim3(b), 0, s, L, F, C, use_b1, use_b2);
}

void launch_daf_k(hipStream_t s, const Level& L, const Fields& F, const Consts& C, int n, double* dq_out) {
    const long nl = (long)(L.ni - 2) * (L.nj - 2);
    const int b = 128;
    const int use_b1 = (n == 1 && C.btype[4] != 5);
    const int build_b2 = (n == 1 && C.btype[5] != 6);
    const int use_b2 = (n == 1 && C.btype[5] != 5);
    hipLaunchKernelGGL(k_daf_k, dim3((unsigned)((nl + b - 1) / b)), dim3(b), 0, s, L, F, C, use_b1, build_b2,
                       use_b2, dq_out);
}

void launch_gtsv_batch(hipStream_t s, int n, int nsys, double* dl, double* d, double* du, double* b, int* info) {
    const int bs = 64;
    hipLaunchKernelGGL(k_gtsv_batch, dim3((unsigned)((nsys + bs - 1) / bs)), dim3(bs), 0, s, n, nsys, dl, d, du, b,
                       info);
}

}  // namespace desalm
