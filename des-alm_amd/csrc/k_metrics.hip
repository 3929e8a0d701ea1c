// k_metrics.hip -- grid metrics and Jacobian on one level, dtev on the fine
// level (src/mg_metrics.F90:56-300): central differences of x, y, z in the
// computational coordinates (one-sided on the boundary planes), inverse
// transformation and J = 1 / det, and the SA pseudo time step.
#include "desalm_internal.h"

namespace desalm {

namespace {

// derivative of coordinate array c along direction with stride s at the
// level-local position p (0..nd-1), spacing h (mg_metrics.F90:64-110)
__device__ __forceinline__ double dcoord(const double* c, long g, long s, int p, int nd, double h) {
    if (p == 0) return h * (c[g + s] - c[g]);
    if (p == nd - 1) return h * (c[g] - c[g - s]);
    return (0.5 * h) * (c[g + s] - c[g - s]);
}

__global__ void k_metrics(Level L, Fields F, Consts C, int fine) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= L.n) return;
    const int ni = L.ni, nj = L.nj, nk = L.nk;
    const int i = (int)(t % ni), j = (int)((t / ni) % nj), k = (int)(t / ((long)ni * nj));
    const long g = L.off + t;
    const long sJ = ni, sK = (long)ni * nj;
    const double* X = F.x[0];
    const double* Y = F.x[1];
    const double* Z = F.x[2];
    const double xc = dcoord(X, g, 1, i, ni, L.dc), yc = dcoord(Y, g, 1, i, ni, L.dc), zc = dcoord(Z, g, 1, i, ni, L.dc);
    const double xe = dcoord(X, g, sJ, j, nj, L.de), ye = dcoord(Y, g, sJ, j, nj, L.de), ze = dcoord(Z, g, sJ, j, nj, L.de);
    const double xz = dcoord(X, g, sK, k, nk, L.dz), yz = dcoord(Y, g, sK, k, nk, L.dz), zz = dcoord(Z, g, sK, k, nk, L.dz);
    const double gj = xc * (ye * zz - yz * ze) + xe * (yz * zc - yc * zz) + xz * (yc * ze - ye * zc);
    const double a = 1.0 / gj;
    F.aj[g] = a;
    F.csi[0][g] = a * (ye * zz - yz * ze);
    F.csi[1][g] = a * (xz * ze - xe * zz);
    F.csi[2][g] = a * (xe * yz - xz * ye);
    F.eta[0][g] = a * (yz * zc - yc * zz);
    F.eta[1][g] = a * (xc * zz - xz * zc);
    F.eta[2][g] = a * (xz * yc - xc * yz);
    F.zet[0][g] = a * (yc * ze - ye * zc);
    F.zet[1][g] = a * (xe * zc - xc * ze);
    F.zet[2][g] = a * (xc * ye - xe * yc);
    if (fine) {
        const double Sc = sqrt(xc * xc + yc * yc + zc * zc);
        const double Se = sqrt(xe * xe + ye * ye + ze * ze);
        const double Sz = sqrt(xz * xz + yz * yz + zz * zz);
        const double SL = fmin(fmin(Sc, Se), Sz);
        double d = SL * fmin(C.cfl2, C.vnn2 * C.ren * SL);
        if (C.unsteady) d = fmin(d, C.delti * 2.0 / 3.0);
        F.dtev[g] = d;
    }
}

}  // namespace

void launch_metrics(hipStream_t s, const Level& L, const Fields& F, const Consts& C, int fine) {
    const int b = 256;
    hipLaunchKernelGGL(k_metrics, dim3((unsigned)((L.n + b - 1) / b)), dim3(b), 0, s, L, F, C, fine);
}

}  // namespace desalm
