// k_alm.hip -- actuator-line model: rhs_lines (src/rhs_lines.F90:47-483)
// and the hub zeroing of bcond_fm (src/bcond_fm.F90:328-355).
//
//   k_alm_elements  one wave per (blade, element): velocity at the
//                   element, angle of attack, polar lookup, lift and drag ->
//                   the element's position and force (forces_lines)
//   k_alm_spread    per spreading node: Gaussian sum over all elements of
//                   its rotor, added to rh(2:4) / J
//   k_alm_hub       one thread per disk node: u = 0 on hub nodes,
//                   i = itur1..itur2
// The tables hold fine-level node indices.  rhs_lines runs on every level
// and tests them against the current level's interior range and arrays,
// as the reference does; the kernels take the current level.  Operand
// order follows the reference routine.
#include "desalm_internal.h"

namespace desalm {

namespace {

constexpr double kPi = 3.141592653589793;
constexpr double kRho = 1178.62;  // rhs_lines.F90:86

// Cl, Cd from polars (rhs_lines.F90:232-265).  The reference walks the
// pairs (a1, a2) = (angle(k), angle(k+1)) from the first and stops at the
// first pair with a1 < aoa <= a2, or at the first a1 >= aoa.  Either way the
// walk ends at b, the first index with aoa <= angle(b): b = 0 gives the first
// values, 0 < b < n the linear interpolation on (b-1, b), no such b the last
// values.  The 64 lanes of the wave find b together, 256 angles of each polar
// per step with all their loads issued first (every lane calls this with the
// same polars and aoa); the two polars of an element are searched at once.
constexpr int kPolarStep = 256;

__device__ void polar_lookup2(const AlmPolars& P, const int (&p)[2], double aoa, double (&cl)[2], double (&cd)[2]) {
    const int lane = (int)(threadIdx.x % 64);
    int o[2], n[2], b[2];
    bool found[2];
#pragma unroll
    for (int s = 0; s < 2; s++) {
        o[s] = P.off[p[s]];
        n[s] = P.cnt[p[s]];
        b[s] = n[s];
        found[s] = false;
    }
    for (int k0 = 0; k0 < max(n[0], n[1]) && !(found[0] && found[1]); k0 += kPolarStep) {
        double v[2][kPolarStep / 64];
#pragma unroll
        for (int s = 0; s < 2; s++)
#pragma unroll
            for (int u = 0; u < kPolarStep / 64; u++) {
                const int k = k0 + 64 * u + lane;
                v[s][u] = !found[s] && k < n[s] ? P.angle[o[s] + k] : 0.0;
            }
#pragma unroll
        for (int s = 0; s < 2; s++)
#pragma unroll
            for (int u = 0; u < kPolarStep / 64; u++) {
                const int k = k0 + 64 * u + lane;
                const unsigned long long hit = __ballot(!found[s] && k < n[s] && aoa <= v[s][u]);
                if (hit) {
                    b[s] = k0 + 64 * u + __ffsll((long long)hit) - 1;
                    found[s] = true;
                }
            }
    }
    // the operands of all three cases, then the case of b
    double a1[2], a2[2], l1[2], l2[2], d1[2], d2[2];
#pragma unroll
    for (int s = 0; s < 2; s++) {
        const int hi = min(max(b[s], 0), n[s] - 1), lo = max(b[s] - 1, 0);
        a1[s] = P.angle[o[s] + lo]; a2[s] = P.angle[o[s] + hi];
        l1[s] = P.cl[o[s] + lo]; l2[s] = P.cl[o[s] + hi];
        d1[s] = P.cd[o[s] + lo]; d2[s] = P.cd[o[s] + hi];
    }
#pragma unroll
    for (int s = 0; s < 2; s++) {
        if (b[s] == 0) {
            cl[s] = l2[s];
            cd[s] = d2[s];
        } else if (b[s] < n[s]) {
            double m = (l2[s] - l1[s]) / (a2[s] - a1[s]);
            cl[s] = m * (aoa - a1[s]) + l1[s];
            m = (d2[s] - d1[s]) / (a2[s] - a1[s]);
            cd[s] = m * (aoa - a1[s]) + d1[s];
        } else {
            cl[s] = l2[s];
            cd[s] = d2[s];
        }
    }
}

__global__ void k_alm_elements(Level L, Fields F, AlmTurbine T, AlmPolars P, double length_scale, double vel_infty,
                               double ren, int jmg, int kmg) {
    // one wave per (blade, element); every lane computes the element, lane 0
    // writes it
    const int e = (int)blockIdx.x;
    if (e >= T.nblades * T.nelems) return;
    const bool writer = threadIdx.x == 0;
    const int bld = 1 + e / T.nelems, elem = 1 + e % T.nelems;
    double* f = T.forces + 6L * e;
    if (writer)
#pragma unroll
        for (int c = 0; c < 6; c++) f[c] = 0.0;  // forces_lines = 0 (:95)
    const double vel_scale = ren * 10E-7 / length_scale;
    const double omega = (vel_infty * T.tsr) / T.tur_rad;
    double angle_aux = T.angle + (bld - 1) * 2 * kPi / T.nblades;
    if (angle_aux > 2 * kPi) angle_aux = angle_aux - 2 * kPi;
    const int nang = (int)lround(angle_aux / T.dtheta) + 1;
    const long ne = T.nelems, na = T.nangles;
    const long ix = (elem - 1) + ne * (nang - 1);
    const int i = T.node_idx[ix], j = T.node_idx[ix + ne * na], k = T.node_idx[ix + 2 * ne * na];
    if (!(i >= 2 && i <= L.ni - 1 && j >= 2 && j <= L.nj - 1 && k >= 2 && k <= L.nk - 1)) return;
    auto np = [&](int c) { return T.node_pos[ix + ne * na * (c - 1)]; };
    const double xf = np(1), yf = np(2), zf = np(3);
    const double x000 = np(4), y000 = np(5), z000 = np(6), x100 = np(7), y100 = np(8);
    const double x110 = np(10), y110 = np(11), x001 = np(16), z001 = np(18);
    const double x101 = np(19), y101 = np(20), y111 = np(23);
    (void)y000;
    (void)x110;
    const long jk = (j - 1) + (long)jmg * (k - 1), ck = (long)jmg * kmg;
    const double radius = T.disk_dat[jk] * length_scale;
    const double chord = T.disk_dat[jk + 4 * ck], twist = T.disk_dat[jk + 5 * ck], inter = T.disk_dat[jk + 6 * ck];
    const long sJ = L.ni, sK = (long)L.ni * L.nj;
    const long g = L.off + (i - 1) + sJ * (j - 1) + sK * (k - 1);
    double uvw[3];
#pragma unroll
    for (int m = 1; m <= 3; m++) {
        const double* q = F.q[m];
        const double down = 1 / ((x100 - x000) * (y110 - y100)) *
                            ((x100 - xf) * ((y110 - yf) * (q[g] + q[g + sJ])) +
                             (xf - x000) * ((yf - y100) * (q[g + 1] + q[g + 1 + sJ])));
        const double up = 1 / ((x101 - x001) * (y111 - y101)) *
                          ((x101 - xf) * ((y111 - yf) * (q[g + sK] + q[g + sJ + sK])) +
                           (xf - x001) * ((yf - y101) * (q[g + 1 + sK] + q[g + 1 + sJ + sK])));
        uvw[m - 1] = down + (up - down) / (z001 - z000) * (zf - z000);
    }
    const double u_x = uvw[0] * vel_scale, u_y = uvw[1] * vel_scale, u_z = uvw[2] * vel_scale;
    const double u_theta = omega * radius + u_y * sin(angle_aux) - u_z * cos(angle_aux);
    const double phi = atan2(u_x, u_theta);
    const double aoa = phi * 180.0 / kPi - twist;
    double clv[2], cdv[2];
    polar_lookup2(P, {T.pol_b[jk], T.pol_a[jk]}, aoa, clv, cdv);
    const double cl_b = clv[0], cd_b = cdv[0], cl_a = clv[1], cd_a = cdv[1];
    const double cl = (cl_a - cl_b) * inter + cl_b;
    const double cd = (cd_a - cd_b) * inter + cd_b;
    const double urel = sqrt(u_theta * u_theta + u_x * u_x);
    const double fL = 0.5 * cl * kRho * (urel * urel) * chord;
    const double fD = 0.5 * cd * kRho * (urel * urel) * chord;
    if (!writer) return;
    f[0] = xf;
    f[1] = yf;
    f[2] = zf;
    f[3] = fL * cos(phi) + fD * sin(phi);
    f[4] = fL * sin(angle_aux) * sin(phi) - fD * sin(angle_aux) * sin(phi);
    f[5] = fL * cos(angle_aux) * sin(phi) - fD * cos(angle_aux) * cos(phi);
}

// Spreading: a workgroup of kSpreadWaves waves owns 64 spreading nodes (one
// per lane) and walks the elements in chunks of 64.  Per chunk the chunk's
// element positions and forces go to LDS, wave w evaluates the Gaussian
// kernel of every element e = w (mod kSpreadWaves) of the chunk at its lane's
// node into LDS, and after a barrier wave 0 adds the chunk's terms to its
// node's sums in element order, as the reference's element loop does.
constexpr int kSpreadWaves = 4;
constexpr int kSpreadChunk = 64;

__global__ void __launch_bounds__(64 * kSpreadWaves) k_alm_spread(Level L, Fields F, AlmTurbine T,
                                                                  double length_scale, double ren) {
    __shared__ double s_f[kSpreadChunk * 6];
    __shared__ double s_k[kSpreadChunk][65];  // [element of the chunk][node]; odd stride: no bank conflicts
    const int ne = T.nblades * T.nelems;
    const int lane = (int)(threadIdx.x % 64), w = (int)(threadIdx.x / 64);
    const int c = (int)blockIdx.x * 64 + lane;
    bool ok = c < T.n_ker;
    int ik = 0, jk = 0, kk = 0;
    if (ok) {
        ik = T.ker_nodes[3 * c];
        jk = T.ker_nodes[3 * c + 1];
        kk = T.ker_nodes[3 * c + 2];
        ok = ik >= 2 && ik <= L.ni - 1 && jk >= 2 && jk <= L.nj - 1 && kk >= 2 && kk <= L.nk - 1;
    }
    const double Ls = length_scale;
    double dV = 0.0, xk = 0.0, yk = 0.0, zk = 0.0;
    if (ok) {
        const double radius = T.ker_dat[3 * c] * Ls;
        ok = radius > T.hub_rad;
        dV = T.ker_dat[3 * c + 2] * (Ls * Ls * Ls);
        xk = T.ker_pos[3 * c];
        yk = T.ker_pos[3 * c + 1];
        zk = T.ker_pos[3 * c + 2];
    }
    const double deltar = fabs(T.tur_rad - T.hub_rad) / T.nelems;
    const double eps = 2 * pow(dV, 1.0 / 3.0);
    const double kscale = 1 / ((eps * eps * eps) * pow(kPi, 3.0 / 2.0));
    double fxa = 0.0, fya = 0.0, fza = 0.0;
    for (int e0 = 0; e0 < ne; e0 += kSpreadChunk) {
        const int n = min(kSpreadChunk, ne - e0);
        for (int v = (int)threadIdx.x; v < 6 * n; v += (int)blockDim.x) s_f[v] = T.forces[6L * e0 + v];
        __syncthreads();
        if (ok)
            for (int e = w; e < n; e += kSpreadWaves) {
                const double* f = s_f + 6 * e;
                const double rdis =
                    sqrt((xk - f[0]) * (xk - f[0]) + (yk - f[1]) * (yk - f[1]) + (zk - f[2]) * (zk - f[2])) * Ls;
                s_k[e][lane] = kscale * exp(-(rdis * rdis) / (eps * eps));
            }
        __syncthreads();
        if (ok && w == 0)
            for (int e = 0; e < n; e++) {
                const double* f = s_f + 6 * e;
                const double ker = s_k[e][lane];
                fxa = fxa + ker * f[3] * deltar;
                fya = fya + ker * f[4] * deltar;
                fza = fza + ker * f[5] * deltar;
            }
        __syncthreads();
    }
    if (!ok || w != 0) return;
    const double vel_scale = ren * 10E-7 / Ls;
    const double s = Ls / (kRho * (vel_scale * vel_scale));
    const long g = L.off + (ik - 1) + (long)L.ni * ((jk - 1) + (long)L.nj * (kk - 1));
    const double rj = 1.0 / F.aj[g];
    F.rh[1][g] = F.rh[1][g] + fxa * s * rj;
    F.rh[2][g] = F.rh[2][g] + fya * s * rj;
    F.rh[3][g] = F.rh[3][g] + fza * s * rj;
}

__global__ void k_alm_hub(Level L, Fields F, AlmTurbine T, double length_scale, int jmg) {
    const int c = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (c >= T.n_disk) return;
    const int di = T.disk_nodes[3 * c], dj = T.disk_nodes[3 * c + 1], dk = T.disk_nodes[3 * c + 2];
    if (!(di >= 1 && di <= L.ni && dj >= 1 && dj <= L.nj && dk >= 1 && dk <= L.nk)) return;
    const double radius = T.disk_dat[(dj - 1) + (long)jmg * (dk - 1)] * length_scale;
    if (!(radius <= T.hub_rad)) return;
    for (int it = T.itur1; it <= T.itur2; it++)
        F.q[1][L.off + (it - 1) + (long)L.ni * ((dj - 1) + (long)L.nj * (dk - 1))] = 0.0;
}

}  // namespace

void launch_alm_rhs(hipStream_t s, const Level& L, const Fields& F, const Alm& A, double ren) {
    for (const AlmTurbine& T : A.t) {
        const int ne = T.nblades * T.nelems;
        hipLaunchKernelGGL(k_alm_elements, dim3(ne), dim3(64), 0, s, L, F, T, A.pol, A.length_scale, A.vel_infty,
                           ren, A.jmg, A.kmg);
        if (T.n_ker > 0)
            hipLaunchKernelGGL(k_alm_spread, dim3((T.n_ker + 63) / 64), dim3(64 * kSpreadWaves), 0, s, L, F, T,
                               A.length_scale, ren);
    }
}

void launch_alm_hub(hipStream_t s, const Level& L, const Fields& F, const Alm& A) {
    for (const AlmTurbine& T : A.t)
        if (T.n_disk > 0)
            hipLaunchKernelGGL(k_alm_hub, dim3((T.n_disk + 127) / 128), dim3(128), 0, s, L, F, T, A.length_scale,
                               A.jmg);
}

}  // namespace desalm
