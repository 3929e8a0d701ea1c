// k_rfg.hip -- random-flow-generator inflow (the method of rfg.F90 and
// rfg_inlet.F90:107-173: Smirnov, Shi & Celik, J. Fluids Eng. 123 (2001)
// 359-371).
//
// At physical time t, every interior node (j, k) of the inflow plane i = 1
// gets u, v, w = um(j, k) + m(j, k) (c(j, k) o v1), where
//   v1 = sqrt(2 / N) sum_n [ p_n cos(theta_n) + q_n sin(theta_n) ],
//   theta_n = sum_d k_nd (ls / ts) / c_d x_d / ls + omega_n t / ts,
// (k_n, omega_n, p_n, q_n) are the N random modes and m, c the eigenvectors
// and the square roots of the eigenvalues of the node's Reynolds-stress
// tensor (rfg.F90:329-374).  One thread per node; the mode table is read by
// every thread (it stays in the caches).
#include "desalm_internal.h"

namespace desalm {

namespace {

__global__ void k_rfg_inflow(Level L, Fields F, Rfg R, double t) {
    const int nj = L.nj, nk = L.nk;
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)(nj - 2) * (nk - 2)) return;
    const int j = 1 + (int)(e % (nj - 2)), k = 1 + (int)(e / (nj - 2));
    const long jk = j + (long)nj * k;              // (j, k) index of the plane tables
    const long g = L.off + (long)L.ni * jk;         // node (i = 0, j, k)
    const double ts = R.ts[jk], ls = R.ls[jk];
    const double c[3] = {R.c[3 * jk], R.c[3 * jk + 1], R.c[3 * jk + 2]};
    const double t1 = t / ts;
    const double x1[3] = {F.x[0][g] / ls, F.x[1][g] / ls, F.x[2][g] / ls};
    const double ups = ls / ts;
    double s[3] = {0.0, 0.0, 0.0};
    for (int n = 0; n < R.n; n++) {
        const double* kk = R.k + 3 * n;
        double th = 0.0;
#pragma unroll
        for (int d = 0; d < 3; d++) th = th + kk[d] * ups / c[d] * x1[d];
        th = th + R.omega[n] * t1;
        const double cs = cos(th), sn = sin(th);
#pragma unroll
        for (int d = 0; d < 3; d++) s[d] = s[d] + R.p[3 * n + d] * cs + R.q[3 * n + d] * sn;
    }
    const double cst = sqrt(2.0 / (double)R.n);
    double v1[3];
#pragma unroll
    for (int d = 0; d < 3; d++) v1[d] = c[d] * (cst * s[d]);
    const double* m = R.m + 9 * jk;  // column-major eigenvectors: m[r + 3 col]
#pragma unroll
    for (int r = 0; r < 3; r++) {
        const double v = m[r] * v1[0] + m[r + 3] * v1[1] + m[r + 6] * v1[2];
        F.q[r + 1][g] = R.um[3 * jk + r] + v;
    }
}

}  // namespace

void launch_rfg_inflow(hipStream_t s, const Level& L, const Fields& F, const Rfg& R, double t) {
    const long n = (long)(L.nj - 2) * (L.nk - 2);
    hipLaunchKernelGGL(k_rfg_inflow, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, s, L, F, R, t);
}

}  // namespace desalm
