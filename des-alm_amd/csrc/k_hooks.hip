// k_hooks.hip -- the two case hooks the reference's mg_driver calls on every
// run of zone 1 on one rank (switched on per context by desalm_set_hooks):
//
//  * conservation (src/conservation.F90:1-174, called at
//    src/mg_driver.F90:64-75 after the time shift): the contravariant flux
//    U/J = (csi . u) / J of the whole i = 1 plane and of the whole i = img
//    plane are summed; on the outlet plane U/J is scaled by fluxin / fluxout
//    and u, v, w are rebuilt from U/J, V/J, W/J and the grid derivatives
//    (x_csi, x_eta, x_zet, ...) of src/mg_metrics.F90:64-200.
//
//  * wf_eta (src/wf_mpi.F90:246-413, called on both j walls after every
//    fine-level solver_daf, src/mg_driver.F90:218-224): for every interior
//    (i, k) of the wall the friction velocity u_tau solves
//      f(u) = (ln(Re y u) / kappa + B) u - u0   if Re y u > 11.3,
//      f(u) = Re y u^2 - u0                      otherwise
//    (u0: wall-parallel speed two nodes off the wall, y: its wall distance;
//    f = 0 where |u Re y - 11.3| < 1e-4), and the first node off the wall
//    gets the log-law (or linear) speed in the direction of the third node's
//    velocity, with nu~ = kappa y+ / Re.  The reference solves for u_tau with
//    a Newton chain that starts each node from the previous node's result
//    and stops at |f| < 1e-4 (newt, src/wf_mpi.F90:634-720); here every node
//    is solved on its own to rounding (wall_root), so the two differ by the
//    reference's stopping tolerance.
#include "desalm_internal.h"

namespace desalm {

namespace {

constexpr int kConsThreads = 1024;

__device__ __forceinline__ double contra(const double* const m[3], const Fields& F, long g) {
    const double tmp = 1.0 / F.aj[g];
    return tmp * (m[0][g] * F.q[1][g] + m[1][g] * F.q[2][g] + m[2][g] * F.q[3][g]);
}

// derivative of a coordinate along stride s at position p of nd, spacing h
// (the same one-sided / central forms as the metrics, mg_metrics.F90:64-110)
__device__ __forceinline__ double dgrid(const double* c, long g, long s, int p, int nd, double h) {
    if (p == 0) return h * (c[g + s] - c[g]);
    if (p == nd - 1) return h * (c[g] - c[g - s]);
    return (0.5 * h) * (c[g + s] - c[g - s]);
}

// One workgroup: both plane sums by a fixed-order tree (deterministic), then
// the outlet update.  Plane node p = j + nj k.
__global__ __launch_bounds__(kConsThreads) void k_conservation(Level L, Fields F) {
    __shared__ double s_in[kConsThreads], s_out[kConsThreads];
    const int ni = L.ni, nj = L.nj, nk = L.nk;
    const long np = (long)nj * nk;
    const int t = threadIdx.x;
    double fin = 0.0, fout = 0.0;
    for (long p = t; p < np; p += kConsThreads) {
        const long g0 = L.off + (long)ni * p;
        fin += contra(F.csi, F, g0);
        fout += contra(F.csi, F, g0 + ni - 1);
    }
    s_in[t] = fin;
    s_out[t] = fout;
    __syncthreads();
    for (int h = kConsThreads / 2; h > 0; h >>= 1) {
        if (t < h) {
            s_in[t] += s_in[t + h];
            s_out[t] += s_out[t + h];
        }
        __syncthreads();
    }
    const double fluxin = s_in[0], fluxout = s_out[0];
    const double delta_e = fluxin - fluxout;
    const long sJ = ni, sK = (long)ni * nj;
    for (long p = t; p < np; p += kConsThreads) {
        const int j = (int)(p % nj), k = (int)(p / nj);
        const long g = L.off + (long)ni * p + ni - 1;
        double u1 = contra(F.csi, F, g);
        const double u2 = contra(F.eta, F, g), u3 = contra(F.zet, F, g);
        u1 = u1 + delta_e * u1 / fluxout;
        const double a = F.aj[g];
        double dx[3][3];  // dx[c][d]: derivative of coordinate c along direction d
#pragma unroll
        for (int c = 0; c < 3; c++) {
            dx[c][0] = dgrid(F.x[c], g, 1, ni - 1, ni, L.dc);
            dx[c][1] = dgrid(F.x[c], g, sJ, j, nj, L.de);
            dx[c][2] = dgrid(F.x[c], g, sK, k, nk, L.dz);
        }
#pragma unroll
        for (int c = 0; c < 3; c++) F.q[1 + c][g] = a * (u1 * dx[c][0] + u2 * dx[c][1] + u3 * dx[c][2]);
    }
}

constexpr double kKapa = 0.41, kBlog = 5.45, kYplm = 11.30, kQmin = 1.0e-08;

// u_tau converged to rounding.  f is increasing in u > 0 on both branches,
// convex on the log branch, and jumps up at Re y u = 11.3: the root is
// sqrt(u0 / (Re y)) when that lies on the quadratic branch, the jump point
// when f changes sign across it, and otherwise the limit of Newton's
// iteration started right of the root on the log branch (monotone for a
// convex increasing function).
__device__ double wall_root(double u0, double yp0, double re0) {
    if (!(u0 > 0.0)) return u0 == 0.0 ? 0.0 : u0 - u0;
    const double c = re0 * yp0;
    const double us = sqrt(u0 / c);
    if (c * us <= 11.3) return us;
    const double lo = 11.3 / c;
    if (log(c * lo) / kKapa * lo + kBlog * lo - u0 >= 0.0) return lo;
    double x = 2.0 * lo;
    for (int t = 0; t < 200 && (log(c * x) / kKapa + kBlog) * x - u0 < 0.0; t++) x = 2.0 * x;
    for (int t = 0; t < 100; t++) {
        const double lg = log(c * x);
        const double fx = (lg / kKapa + kBlog) * x - u0;
        const double step = fx / (lg / kKapa + kBlog + 1.0 / kKapa);
        const double xn = x - step;
        if (!(xn < x) || xn <= lo) break;
        x = xn;
    }
    return x;
}

// wf_eta on the wall j = 0 (jb = 1) or j = nj - 1 (jb = 2); one thread per
// interior (i, k).  Reads q two nodes off the wall, writes q one node off.
__global__ void k_wall_eta(Level L, Fields F, double ren, int jb) {
    const int ni = L.ni, nj = L.nj, nk = L.nk;
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)(ni - 2) * (nk - 2)) return;
    const int i = 1 + (int)(e % (ni - 2)), k = 1 + (int)(e / (ni - 2));
    const int j = jb == 1 ? 0 : nj - 1, j1 = jb == 1 ? 1 : nj - 2, j2 = jb == 1 ? 2 : nj - 3;
    const long sJ = ni, sK = (long)ni * nj;
    const long l = L.off + i + sJ * j + sK * k;
    const long l1 = L.off + i + sJ * j1 + sK * k;
    const long l2 = L.off + i + sJ * j2 + sK * k;
    const double u = F.q[1][l2], v = F.q[2][l2], w = F.q[3][l2];
    const double e0 = F.eta[0][l], e1 = F.eta[1][l], e2 = F.eta[2][l];
    const double vmod = sqrt(e0 * e0 + e1 * e1 + e2 * e2);
    const double qnor = (u * e0 + v * e1 + w * e2) / vmod;
    const double qmag2 = u * u + v * v + w * w;
    const double ut = wall_root(sqrt(qmag2 - qnor * qnor), F.wd[l2], ren);
    if (qmag2 > kQmin) {
        const double yplus = F.wd[l1] * ren * ut;
        const double deltb = 0.0;
        const double qw = yplus > kYplm ? (log(yplus) / kKapa + kBlog - deltb) * ut : yplus * ut;
        double xc[3], xz[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            xc[c] = 0.5 * (F.x[c][l + 1] - F.x[c][l - 1]);
            xz[c] = 0.5 * (F.x[c][l + sK] - F.x[c][l - sK]);
        }
        const double sg11 = sqrt(xc[0] * xc[0] + xc[1] * xc[1] + xc[2] * xc[2]);
        const double sg33 = sqrt(xz[0] * xz[0] + xz[1] * xz[1] + xz[2] * xz[2]);
        double qc1 = (F.csi[0][l2] * u + F.csi[1][l2] * v + F.csi[2][l2] * w) * sg11;
        double qc2 = (F.zet[0][l2] * u + F.zet[1][l2] * v + F.zet[2][l2] * w) * sg33;
        const double cgama = (xc[0] * xz[0] + xc[1] * xz[1] + xc[2] * xz[2]) / (sg11 * sg33);
        double fbeta = sqrt(qc1 * qc1 + qc2 * qc2 + 2.0 * qc1 * qc2 * cgama);
        fbeta = (qc1 + qc2 * cgama) / fbeta;
        fbeta = fbeta < 1.0 ? fbeta : 1.0;
        fbeta = fbeta > -1.0 ? fbeta : -1.0;
        fbeta = qc2 > 0.0 ? acos(fbeta) : -acos(fbeta);
        const double qstar = qw;
        qc2 = qstar * sin(fbeta) / sin(acos(cgama));
        qc1 = qstar * cos(fbeta) - qc2 * cgama;
#pragma unroll
        for (int c = 0; c < 3; c++) F.q[1 + c][l1] = qc1 * xc[c] / sg11 + qc2 * xz[c] / sg33;
        F.q[4][l1] = kKapa * yplus / ren;
    } else {
        F.q[1][l1] = 0.0;
        F.q[2][l1] = 0.0;
        F.q[3][l1] = 0.0;
        F.q[4][l1] = 0.1 / ren;
    }
}

}  // namespace

void launch_conservation(hipStream_t s, const Level& L, const Fields& F) {
    hipLaunchKernelGGL(k_conservation, dim3(1), dim3(kConsThreads), 0, s, L, F);
}

void launch_wall_functions(hipStream_t s, const Level& L, const Fields& F, double ren) {
    const long n = (long)(L.ni - 2) * (L.nk - 2);
    const unsigned b = 128, nb = (unsigned)((n + b - 1) / b);
    // the j = 1 wall first: on thin grids the second wall reads what the first wrote
    hipLaunchKernelGGL(k_wall_eta, dim3(nb), dim3(b), 0, s, L, F, ren, 1);
    hipLaunchKernelGGL(k_wall_eta, dim3(nb), dim3(b), 0, s, L, F, ren, 2);
}

}  // namespace desalm
