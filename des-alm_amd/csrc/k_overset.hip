// k_overset.hip -- overset boundary interpolation of the fine level
// (src/mg_bintp_mom.F90:2-172, src/mg_bintp_tur.F90:2-182) for the donor
// table of one zone.
//
// The reference forms every donor value from the current q into a send
// buffer, exchanges it (alltoallv) and only then writes the fringe nodes.
// Here one kernel gathers all donor values into a buffer and a second one
// scatters them, so a fringe node that is also a host corner of another
// donor contributes its pre-interpolation value, as in the reference.
#include "desalm_internal.h"

namespace desalm {

namespace {

// trilinear value of the nc components m0.. of q at donor d
// (mg_bintp_mom.F90:85-131: weights and summation order of the reference)
__global__ void __launch_bounds__(256) k_bintp_gather(Level L, Fields F, Overset O, int m0, int nc) {
    const long d = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= O.n) return;
    const long sJ = L.ni, sK = (long)L.ni * L.nj;
    const long h = L.off + O.host[d];
    const double xco = O.coef[3 * d], yco = O.coef[3 * d + 1], zco = O.coef[3 * d + 2];
    const double one = 1.0;
    const double c000 = (one - xco) * (one - yco) * (one - zco);
    const double c100 = xco * (one - yco) * (one - zco);
    const double c010 = (one - xco) * yco * (one - zco);
    const double c001 = (one - xco) * (one - yco) * zco;
    const double c101 = xco * (one - yco) * zco;
    const double c011 = (one - xco) * yco * zco;
    const double c110 = xco * yco * (one - zco);
    const double c111 = xco * yco * zco;
    for (int c = 0; c < nc; c++) {
        const double* q = F.q[m0 + c];
        O.buf[d * nc + c] = q[h] * c000 + q[h + 1] * c100 + q[h + sJ] * c010 + q[h + sK] * c001 +
                            q[h + 1 + sK] * c101 + q[h + sJ + sK] * c011 + q[h + 1 + sJ] * c110 +
                            q[h + 1 + sJ + sK] * c111;
    }
}

// q(1:4) at the fringe nodes (mg_bintp_mom.F90:141-163); the value of
// fringe node d sits in buffer slot O.slot[d] (slot d without a slot table)
__global__ void __launch_bounds__(256) k_bintp_scatter_mom(Level L, Fields F, Overset O) {
    const long d = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= O.n) return;
    const long l = L.off + O.node[d];
    const long b = O.slot ? O.slot[d] : d;
#pragma unroll
    for (int m = 0; m < 4; m++) F.q[m][l] = O.buf[4 * b + m];
}

// nu~ at the fringe nodes, then its floor and xnut = nu~ fv1
// (mg_bintp_tur.F90:141-177, cv1^3 = 357.911 at :23)
__global__ void __launch_bounds__(256) k_bintp_scatter_tur(Level L, Fields F, Consts C, Overset O, double cvin,
                                                           int des) {
    const long d = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= O.n) return;
    const long l = L.off + O.node[d];
    double q5 = O.buf[O.slot ? O.slot[d] : d];
    if (des) {
        q5 = fmax(cvin, q5);
        const double chi = q5 / C.rei;
        const double chic = chi * chi * chi;
        const double fv1 = chic / (chic + 357.911);
        F.xnut[l] = q5 * fv1;
    }
    F.q[4][l] = q5;
}

}  // namespace

void launch_bintp_gather(hipStream_t s, const Level& L, const Fields& F, const Overset& O, int tur) {
    if (O.n <= 0) return;
    const unsigned g = (unsigned)((O.n + 255) / 256);
    hipLaunchKernelGGL(k_bintp_gather, dim3(g), dim3(256), 0, s, L, F, O, tur ? 4 : 0, tur ? 1 : 4);
}

void launch_bintp_scatter(hipStream_t s, const Level& L, const Fields& F, const Consts& C, const Overset& O, int tur,
                          double cvin, int des) {
    if (O.n <= 0) return;
    const unsigned g = (unsigned)((O.n + 255) / 256);
    if (!tur) hipLaunchKernelGGL(k_bintp_scatter_mom, dim3(g), dim3(256), 0, s, L, F, O);
    else hipLaunchKernelGGL(k_bintp_scatter_tur, dim3(g), dim3(256), 0, s, L, F, C, O, cvin, des);
}

void launch_bintp(hipStream_t s, const Level& L, const Fields& F, const Consts& C, const Overset& O, int tur,
                  double cvin, int des) {
    launch_bintp_gather(s, L, F, O, tur);
    launch_bintp_scatter(s, L, F, C, O, tur, cvin, des);
}

}  // namespace desalm
