// k_sa.hip -- Spalart-Allmaras / DES eddy-viscosity update on the fine level.
//
// des_eddy (src/des_eddy.F90:132-297): vorticity measure once, then itk
// sub-iterations of explicit right-hand side -> three ADI line solves ->
// relaxed update -> boundary conditions, clamp and xnut = nu~ fv1.
// Kernels:
//   k_sa_prep   U/J (stored, rhs_contra_j) and wo (rhs_sa_source)
//   k_sa_rhs    rhs_sa_grdnu, rhs_sa_prod, rhs_sa_viscous, rhs_sa_flux,
//               rhs_sa_unst_visc, rhs_sa_implicit_source, boundary zeroing,
//               and the ADI operators' face coefficients and dtev J / 2
//   k_sa_adi_i  rhs_sa_adi_solver sweep in i (rows staged through LDS)
//   k_sa_adi<D> the sweeps in j and k (D = 2 also applies the update of
//               des_eddy.F90:260-272)
//   k_sa_bc<D>  rhs_sa_bcond faces, k_sa_clamp its nu~ floor and xnut
#include "desalm_internal.h"
#include "line_tile.h"

namespace desalm {

namespace {

// Spalart-Allmaras constants (des_eddy.F90:63-73)
constexpr double cb1 = 0.1355, cb2 = 0.622, sig = 2.0 / 3.0, capa = 0.41, cw1 = 3.23906782, cw2 = 0.3,
                 cw3h = 64.0, cv1c = 357.911, cv2 = 5.0;

__global__ void k_sa_prep(Level L, Fields F) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= L.n) return;
    const int ni = L.ni, nj = L.nj, nk = L.nk;
    const int i = (int)(t % ni), j = (int)((t / ni) % nj), k = (int)(t / ((long)ni * nj));
    const long g = L.off + t;
    const double u = F.q[1][g], v = F.q[2][g], w = F.q[3][g];
    const double tmp = 1.0 / F.aj[g];
    F.ucn[0][t] = tmp * (F.csi[0][g] * u + F.csi[1][g] * v + F.csi[2][g] * w);
    F.ucn[1][t] = tmp * (F.eta[0][g] * u + F.eta[1][g] * v + F.eta[2][g] * w);
    F.ucn[2][t] = tmp * (F.zet[0][g] * u + F.zet[1][g] * v + F.zet[2][g] * w);
    if (i == 0 || i == ni - 1 || j == 0 || j == nj - 1 || k == 0 || k == nk - 1) {
        F.wo[t] = 0.0;
        return;
    }
    // rhs_sa_source.F90:44-93 (interior only)
    const long sI = 1, sJ = ni, sK = (long)ni * nj;
    const double dc2 = 0.5 * L.dc, de2 = 0.5 * L.de, dz2 = 0.5 * L.dz;
    const double* U = F.q[1];
    const double* V = F.q[2];
    const double* W = F.q[3];
    const double uc = dc2 * (U[g + sI] - U[g - sI]), vc = dc2 * (V[g + sI] - V[g - sI]), wc = dc2 * (W[g + sI] - W[g - sI]);
    const double ue = de2 * (U[g + sJ] - U[g - sJ]), ve = de2 * (V[g + sJ] - V[g - sJ]), we = de2 * (W[g + sJ] - W[g - sJ]);
    const double uz = dz2 * (U[g + sK] - U[g - sK]), vz = dz2 * (V[g + sK] - V[g - sK]), wz = dz2 * (W[g + sK] - W[g - sK]);
    const double u2 = uc * F.csi[1][g] + ue * F.eta[1][g] + uz * F.zet[1][g];
    const double u3 = uc * F.csi[2][g] + ue * F.eta[2][g] + uz * F.zet[2][g];
    const double v1 = vc * F.csi[0][g] + ve * F.eta[0][g] + vz * F.zet[0][g];
    const double v3 = vc * F.csi[2][g] + ve * F.eta[2][g] + vz * F.zet[2][g];
    const double w1 = wc * F.csi[0][g] + we * F.eta[0][g] + wz * F.zet[0][g];
    const double w2 = wc * F.csi[1][g] + we * F.eta[1][g] + wz * F.zet[1][g];
    const double w12 = 0.5 * (u2 - v1), w13 = 0.5 * (u3 - w1), w23 = 0.5 * (v3 - w2);
    F.wo[t] = 2.0 * (w12 * w12 + w13 * w13 + w23 * w23);
}

// rhs_sa_viscous face flux between level-local node t and t + s_D
template <int D>
__device__ __forceinline__ double sa_visc_face(const Level& L, const Fields& F, const Consts& C, long t, long sI,
                                               long sJ, long sK) {
    const long sD = D == 0 ? sI : D == 1 ? sJ : sK;
    const long g0 = L.off + t, g1 = g0 + sD;
    const double* q5 = F.q[4];
    const double hc = L.dc, he = L.de, hz = L.dz;
    const double hc4 = 0.25 * hc, he4 = 0.25 * he, hz4 = 0.25 * hz;
    double vcs, vet, vze;
    if (D == 0) {
        vcs = hc * (q5[g1] - q5[g0]);
        vet = he4 * (q5[g1 + sJ] - q5[g1 - sJ] + q5[g0 + sJ] - q5[g0 - sJ]);
        vze = hz4 * (q5[g1 + sK] - q5[g1 - sK] + q5[g0 + sK] - q5[g0 - sK]);
    } else if (D == 1) {
        vcs = hc4 * (q5[g1 + sI] - q5[g1 - sI] + q5[g0 + sI] - q5[g0 - sI]);
        vet = he * (q5[g1] - q5[g0]);
        vze = hz4 * (q5[g1 + sK] - q5[g1 - sK] + q5[g0 + sK] - q5[g0 - sK]);
    } else {
        vcs = hc4 * (q5[g1 + sI] - q5[g1 - sI] + q5[g0 + sI] - q5[g0 - sI]);
        vet = he4 * (q5[g1 + sJ] - q5[g1 - sJ] + q5[g0 + sJ] - q5[g0 - sJ]);
        vze = hz * (q5[g1] - q5[g0]);
    }
    const double a1 = F.csi[0][g1] + F.csi[0][g0], a2 = F.csi[1][g1] + F.csi[1][g0], a3 = F.csi[2][g1] + F.csi[2][g0];
    const double b1 = F.eta[0][g1] + F.eta[0][g0], b2 = F.eta[1][g1] + F.eta[1][g0], b3 = F.eta[2][g1] + F.eta[2][g0];
    const double c1 = F.zet[0][g1] + F.zet[0][g0], c2 = F.zet[1][g1] + F.zet[1][g0], c3 = F.zet[2][g1] + F.zet[2][g0];
    const double f1 = D == 0 ? a1 : D == 1 ? b1 : c1;
    const double f2 = D == 0 ? a2 : D == 1 ? b2 : c2;
    const double f3 = D == 0 ? a3 : D == 1 ? b3 : c3;
    const double cc1 = f1 * a1 + f2 * a2 + f3 * a3;
    const double cc2 = f1 * b1 + f2 * b2 + f3 * b3;
    const double cc3 = f1 * c1 + f2 * c2 + f3 * c3;
    const double ret1 = (C.rei + q5[g1]) / sig, ret0 = (C.rei + q5[g0]) / sig;
    const double vtj = 0.25 * (ret1 + ret0) / (F.aj[g1] + F.aj[g0]);
    return vtj * (cc1 * vcs + cc2 * vet + cc3 * vze);
}

// rhs_sa_flux face flux (first-order upwind FDS, rhs_sa_flux.F90:226-294)
template <int D>
__device__ __forceinline__ double sa_conv_face(const Level& L, const Fields& F, long t, long sD) {
    const double u1 = F.ucn[D][t + sD], u0 = F.ucn[D][t];
    const double q1 = F.q[4][L.off + t + sD], q0 = F.q[4][L.off + t];
    const double eigh = 0.5 * fabs(u1 + u0);
    return 0.5 * (q1 * u1 + q0 * u0 - eigh * (q1 - q0));
}

// Face and node coefficients of the three ADI operators
// (rhs_sa_adi_solver.F90:53-75 and copies), for every fine node:
//   av_D = |U/J face average| + (1/Re + nu~ face average)/sig * |grad csi_D|^2_face / (J + J')
// on the face between the node and its +D neighbour, and dtj = dtev J / 2.
// nu~ does not change between the three sweeps of one sub-iteration, so
// k_sa_rhs computes them while it holds the node's metrics.
__device__ __forceinline__ void sa_adi_coef(const Level& L, const Fields& F, const Consts& C, long t, int i, int j,
                                            int k) {
    const int ni = L.ni, nj = L.nj, nk = L.nk;
    const long o = L.off;
    const double* q5 = F.q[4];
    F.dtj[t] = 0.5 * F.dtev[o + t] * F.aj[o + t];
    const int pos[3] = {i, j, k};
    const int nd[3] = {ni, nj, nk};
    const long st[3] = {1, (long)ni, (long)ni * nj};
#pragma unroll
    for (int D = 0; D < 3; D++) {
        if (pos[D] >= nd[D] - 1) continue;
        const double* m0 = D == 0 ? F.csi[0] : D == 1 ? F.eta[0] : F.zet[0];
        const double* m1 = D == 0 ? F.csi[1] : D == 1 ? F.eta[1] : F.zet[1];
        const double* m2 = D == 0 ? F.csi[2] : D == 1 ? F.eta[2] : F.zet[2];
        const long p = t, p1 = t + st[D];
        const double gx = m0[o + p1] + m0[o + p], gy = m1[o + p1] + m1[o + p], gz = m2[o + p1] + m2[o + p];
        const double ev = (C.rei + 0.5 * (q5[o + p1] + q5[o + p])) / sig;
        const double ucna = 0.5 * fabs(F.ucn[D][p1] + F.ucn[D][p]);
        F.av[D][t] = ucna + ev * (gx * gx + gy * gy + gz * gz) / (F.aj[o + p1] + F.aj[o + p]);
    }
}

__global__ void __launch_bounds__(256) k_sa_rhs(Level L, Fields F, Consts C) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= L.n) return;
    const int ni = L.ni, nj = L.nj, nk = L.nk;
    const int i = (int)(t % ni), j = (int)((t / ni) % nj), k = (int)(t / ((long)ni * nj));
    sa_adi_coef(L, F, C, t, i, j, k);
    if (i == 0 || i == ni - 1 || j == 0 || j == nj - 1 || k == 0 || k == nk - 1) {
        F.rh5[t] = 0.0;  // des_eddy.F90:231-236
        return;
    }
    const long sI = 1, sJ = ni, sK = (long)ni * nj;
    const long g = L.off + t;
    const double* q5 = F.q[4];
    // gradient of nu~ (rhs_sa_grdnu.F90:22-60)
    double dnu;
    {
        const double dc2 = 0.5 * L.dc, de2 = 0.5 * L.de, dz2 = 0.5 * L.dz;
        const double qip = q5[g + sI] / F.aj[g + sI], qim = q5[g - sI] / F.aj[g - sI];
        const double qjp = q5[g + sJ] / F.aj[g + sJ], qjm = q5[g - sJ] / F.aj[g - sJ];
        const double qkp = q5[g + sK] / F.aj[g + sK], qkm = q5[g - sK] / F.aj[g - sK];
        double d[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            d[c] = dc2 * (qip * F.csi[c][g + sI] - qim * F.csi[c][g - sI]);
            d[c] = d[c] + de2 * (qjp * F.eta[c][g + sJ] - qjm * F.eta[c][g - sJ]);
            d[c] = d[c] + dz2 * (qkp * F.zet[c][g + sK] - qkm * F.zet[c][g - sK]);
        }
        const double aj2 = F.aj[g] * F.aj[g];
        dnu = aj2 * (d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    }
    // production, destruction, cb2 diffusion (rhs_sa_prod.F90:97-165)
    double psa;
    {
        const double wdv = F.wd[g];
        if (wdv == 0.0) {
            psa = 0.0;
        } else {
            const double q = q5[g];
            const double chi = q / C.rei;
            const double chic = chi * chi * chi;
            const double fv1 = chic / (chic + cv1c);
            const double fv2 = pow(1.0 + chi / cv2, -3.0);
            const double fv3 = ((1.0 + chi * fv1) * (1.0 - fv2)) / chi;
            const double vtcm = sqrt(2.0 * F.wo[t]);
            const double cawd = capa * capa * wdv * wdv;
            const double vtc_org = fv3 * vtcm + q * fv2 / cawd;
            const double vtc = fmax(vtc_org, 0.00001);
            psa = cb1 * vtc * q;
            const double rr = q / (vtc * cawd);
            const double rr6 = rr * rr * rr * rr * rr * rr;
            const double grr = fmax(0.0, rr + cw2 * (rr6 - rr));
            const double grr6 = grr * grr * grr * grr * grr * grr;
            const double fw = grr * pow((1.0 + cw3h) / (grr6 + cw3h), 1.0 / 6.0);
            const double vwd = q / wdv;
            psa = psa - cw1 * fw * vwd * vwd;
            psa = psa + cb2 * dnu / sig;
        }
    }
    // diffusion (rhs_sa_viscous.F90:58-188)
    double visc = sa_visc_face<0>(L, F, C, t, sI, sJ, sK) - sa_visc_face<0>(L, F, C, t - sI, sI, sJ, sK);
    visc = visc + sa_visc_face<1>(L, F, C, t, sI, sJ, sK) - sa_visc_face<1>(L, F, C, t - sJ, sI, sJ, sK);
    visc = visc + sa_visc_face<2>(L, F, C, t, sI, sJ, sK) - sa_visc_face<2>(L, F, C, t - sK, sI, sJ, sK);
    // convection (rhs_sa_flux.F90:226-303)
    double r = sa_conv_face<0>(L, F, t, sI) - sa_conv_face<0>(L, F, t - sI, sI);
    r = r + sa_conv_face<1>(L, F, t, sJ) - sa_conv_face<1>(L, F, t - sJ, sJ);
    r = r + sa_conv_face<2>(L, F, t, sK) - sa_conv_face<2>(L, F, t - sK, sK);
    // BDF2, diffusion and source (rhs_sa_unst_visc.F90:188-200)
    const double dtj = 2.0 * C.delti * F.aj[g];
    const double dq5dt = C.e_source * (3.0 * q5[g] - 4.0 * F.qn[4][g] + F.qnm1[4][g]) / dtj;
    r = r + dq5dt - visc - psa / F.aj[g];
    // explicit source form (rhs_sa_implicit_source.F90:244-256, dk = 1)
    F.rh5[t] = -F.dtev[g] * F.aj[g] * r;
}

// one ADI sweep (rhs_sa_adi_solver.F90:44-244), one thread per line:
//   aw = -dtj (U_(s-1) + av_(s-1/2)), ap = 1 + dtj (av_(s-1/2) + av_(s+1/2)),
//   ae = dtj (U_(s+1) - av_(s+1/2)).
// D = 2 also does the relaxed update of des_eddy.F90:260-272:
//   q5 += (-q5 + qold + rh) / (1 + 1.5 e_source dtev / delti)
// rows loaded per group, one group ahead of the elimination (16 measured
// best among 4, 8 and 16 at 129^3)
constexpr int kPF = 16;

template <int D>
__global__ void __launch_bounds__(128) k_sa_adi(Level L, Fields F, Consts C) {
    const int ni = L.ni, nj = L.nj, nk = L.nk;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long base, sD;
    int len;
    if (D == 0) {
        if (t >= (long)(nj - 2) * (nk - 2)) return;
        base = (long)ni * ((1 + t % (nj - 2)) + (long)nj * (1 + t / (nj - 2)));
        sD = 1;
        len = ni - 2;
    } else if (D == 1) {
        if (t >= (long)(ni - 2) * (nk - 2)) return;
        base = (1 + t % (ni - 2)) + (long)ni * nj * (1 + t / (ni - 2));
        sD = ni;
        len = nj - 2;
    } else {
        if (t >= (long)(ni - 2) * (nj - 2)) return;
        base = (1 + t % (ni - 2)) + (long)ni * (1 + t / (ni - 2));
        sD = (long)ni * nj;
        len = nk - 2;
    }
    const double* __restrict__ U = F.ucn[D];
    const double* __restrict__ av = F.av[D];
    const double* __restrict__ dtj = F.dtj;
    double* __restrict__ rh = F.rh5;
    double* __restrict__ cp = F.cp[0];
    double avm = av[base];
    double um = U[base], u0 = U[base + sD];
    double cprev = 0.0, dprev = 0.0, denprev = 1.0;
    unsigned piv = 0;
    double b_av[kPF], b_dt[kPF], b_up[kPF], b_rh[kPF];
#pragma unroll
    for (int u = 0; u < kPF; u++) {
        const long p = base + (1 + u) * sD;
        const bool ok = u < len;
        b_av[u] = ok ? av[p] : 0.0;
        b_dt[u] = ok ? dtj[p] : 0.0;
        b_up[u] = ok ? U[p + sD] : 0.0;
        b_rh[u] = ok ? rh[p] : 0.0;
    }
    for (int s0 = 0; s0 < len; s0 += kPF) {
        double n_av[kPF], n_dt[kPF], n_up[kPF], n_rh[kPF];
#pragma unroll
        for (int u = 0; u < kPF; u++) {
            const int s = s0 + kPF + u;
            const long p = base + (1 + s) * sD;
            const bool ok = s < len;
            n_av[u] = ok ? av[p] : 0.0;
            n_dt[u] = ok ? dtj[p] : 0.0;
            n_up[u] = ok ? U[p + sD] : 0.0;
            n_rh[u] = ok ? rh[p] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kPF; u++) {
            const int s = s0 + u;
            if (s < len) {
                const long p = base + (1 + s) * sD;
                const double avp = b_av[u], dt = b_dt[u], up = b_up[u];
                const double aw = -dt * (um + avm);
                const double ap = 1.0 + dt * (avm + avp);
                const double ae = dt * (up - avp);
                const double aq = b_rh[u];
                double den;
                if (s == 0) {
                    den = ap;
                } else {
                    if (fabs(denprev) < fabs(aw)) piv++;
                    den = ap - aw * cprev;
                }
                const double inv = 1.0 / den;
                cprev = ae * inv;
                dprev = (s == 0) ? aq * inv : (aq - aw * dprev) * inv;
                denprev = den;
                cp[p] = cprev;
                rh[p] = dprev;
                avm = avp;
                um = u0;
                u0 = up;
            }
        }
#pragma unroll
        for (int u = 0; u < kPF; u++) { b_av[u] = n_av[u]; b_dt[u] = n_dt[u]; b_up[u] = n_up[u]; b_rh[u] = n_rh[u]; }
    }
    auto finish = [&](long p, double x) {
        if (D < 2) {
            rh[p] = x;
        } else {
            const long g = L.off + p;
            const double qv = F.q[4][g];
            const double qe5 = qv + x;  // qold == q5 at this point of the sub-iteration
            const double dum = 1.0 + 1.5 * C.e_source * F.dtev[g] / C.delti;
            F.q[4][g] = qv + (-qv + qe5) / dum;
        }
    };
    double x = dprev;
    finish(base + (long)len * sD, x);
    for (int s0 = len - 2; s0 >= 0; s0 -= kPF) {
        double dd[kPF], cc[kPF], qq[kPF], dv[kPF];
#pragma unroll
        for (int u = 0; u < kPF; u++) {
            const int s = s0 - u;
            const long p = base + (1 + s) * sD;
            dd[u] = s >= 0 ? rh[p] : 0.0;
            cc[u] = s >= 0 ? cp[p] : 0.0;
            if (D == 2) {
                qq[u] = s >= 0 ? F.q[4][L.off + p] : 0.0;
                dv[u] = s >= 0 ? F.dtev[L.off + p] : 0.0;
            }
        }
#pragma unroll
        for (int u = 0; u < kPF; u++) {
            const int s = s0 - u;
            if (s >= 0) {
                x = dd[u] - cc[u] * x;
                const long p = base + (1 + s) * sD;
                if (D < 2) {
                    rh[p] = x;
                } else {
                    // the update of finish() with q5 and dtev already loaded
                    const double qv = qq[u];
                    const double qe5 = qv + x;
                    const double dum = 1.0 + 1.5 * C.e_source * dv[u] / C.delti;
                    F.q[4][L.off + p] = qv + (-qv + qe5) / dum;
                }
            }
        }
    }
    if (piv) atomicAdd(F.pivots, (unsigned long long)piv);
}

// The i-direction sweep of k_sa_adi with its rows staged through LDS
// (line_tile.h); one thread per line, kSL lines per workgroup, and a second
// wave that only helps to move the tiles.  Chunks of av,
// U/J, dtj and rh are loaded one chunk ahead of the elimination; c' and d'
// go back to HBM per chunk except for the last one.  Same arithmetic as
// k_sa_adi<0>.
constexpr int kSL = 64;
constexpr int kSNT = 2 * kSL;  // threads per workgroup

__global__ void __launch_bounds__(kSNT) k_sa_adi_i(Level L, Fields F) {
    constexpr int NT = kSNT;
    using IOr = TileIO<NT, kSL, kTR, kTS>;          // rows s0.. at node base + 1 + s
    using IOa = TileIO<NT, kSL, kTR + 1, kTR + 1>;  // av at nodes base + s0 .. base + s0 + kTR
    using IOu = TileIO<NT, kSL, kTR + 2, kTR + 3>;  // U/J at nodes base + s0 .. base + s0 + kTR + 1
    __shared__ long s_base[kSL];
    __shared__ double s_av[kSL * (kTR + 1)];
    __shared__ double s_u[kSL * (kTR + 3)];
    __shared__ double s_dt[kSL * kTS];
    __shared__ double s_r[kSL * kTS];
    __shared__ double s_c[kSL * kTS];

    const long nlines = (long)(L.nj - 2) * (L.nk - 2);
    const long t0 = (long)blockIdx.x * kSL;
    const int nl = (int)min((long)kSL, nlines - t0);
    const int l = (int)threadIdx.x;
    const bool act = l < nl;  // threads kSL.. only move tiles
    if (l < kSL) s_base[l] = act ? iline_base(L, t0 + l) : 0;
    __syncthreads();

    const int len = L.ni - 2;
    const int nch = (len + kTR - 1) / kTR;
    const double* __restrict__ U = F.ucn[0];
    const double* __restrict__ av = F.av[0];
    double* __restrict__ rh = F.rh5;
    double* __restrict__ cp = F.cp[0];
    const int lc = l < kSL ? l : 0;
    double* myr = s_r + lc * kTS;
    double* myc = s_c + lc * kTS;
    const double* mydt = s_dt + lc * kTS;
    const double* mya = s_av + lc * (kTR + 1);
    const double* myu = s_u + lc * (kTR + 3);

    double va[IOa::kPer], vu[IOu::kPer], vt[IOr::kPer], vr[IOr::kPer];
    IOa::load(va, av, s_base, nl, 0, len + 1);
    IOu::load(vu, U, s_base, nl, 0, len + 2);
    IOr::load(vt, F.dtj + 1, s_base, nl, 0, len);
    IOr::load(vr, rh + 1, s_base, nl, 0, len);

    double cprev = 0.0, dprev = 0.0, denprev = 1.0;
    unsigned piv = 0;
    for (int ch = 0; ch < nch; ch++) {
        const int s0 = ch * kTR;
        __syncthreads();
        IOa::to_lds(va, s_av);
        IOu::to_lds(vu, s_u);
        IOr::to_lds(vt, s_dt);
        IOr::to_lds(vr, s_r);
        __syncthreads();
        if (ch + 1 < nch) {
            const int s1 = s0 + kTR;
            IOa::load(va, av, s_base, nl, s1, len + 1);
            IOu::load(vu, U, s_base, nl, s1, len + 2);
            IOr::load(vt, F.dtj + 1, s_base, nl, s1, len);
            IOr::load(vr, rh + 1, s_base, nl, s1, len);
        }
        if (act) {
            for (int u = 0; u < kTR; u++) {
                const int s = s0 + u;
                if (s >= len) break;
                const double um = myu[u], up = myu[u + 2];
                const double avm = mya[u], avp = mya[u + 1], dt = mydt[u];
                const double aw = -dt * (um + avm);
                const double ap = 1.0 + dt * (avm + avp);
                const double ae = dt * (up - avp);
                const double aq = myr[u];
                double den;
                if (s == 0) {
                    den = ap;
                } else {
                    if (fabs(denprev) < fabs(aw)) piv++;
                    den = ap - aw * cprev;
                }
                const double inv = 1.0 / den;
                cprev = ae * inv;
                dprev = (s == 0) ? aq * inv : (aq - aw * dprev) * inv;
                denprev = den;
                myc[u] = cprev;
                myr[u] = dprev;
            }
        }
        __syncthreads();
        if (ch + 1 < nch) {
            IOr::store(s_c, cp + 1, s_base, nl, s0, len);
            IOr::store(s_r, rh + 1, s_base, nl, s0, len);
        }
    }
    double x = 0.0;
    for (int ch = nch - 1; ch >= 0; ch--) {
        const int s0 = ch * kTR;
        if (ch < nch - 1) {
            __syncthreads();
            IOr::to_lds(vt, s_c);
            IOr::to_lds(vr, s_r);
            __syncthreads();
        }
        if (ch > 0) {
            IOr::load(vt, cp + 1, s_base, nl, s0 - kTR, len);
            IOr::load(vr, rh + 1, s_base, nl, s0 - kTR, len);
        }
        if (act) {
            for (int u = kTR - 1; u >= 0; u--) {
                const int s = s0 + u;
                if (s >= len) continue;
                x = (s == len - 1) ? myr[u] : myr[u] - myc[u] * x;
                myr[u] = x;
            }
        }
        __syncthreads();
        IOr::store(s_r, rh + 1, s_base, nl, s0, len);
    }
    if (piv) atomicAdd(F.pivots, (unsigned long long)piv);
}

// rhs_sa_bcond faces (rhs_sa_bcond.F90:40-269); faces of direction D, both sides
template <int D>
__global__ void k_sa_bc(Level L, Fields F, Consts C) {
    const int ni = L.ni, nj = L.nj, nk = L.nk;
    const long sI = 1, sJ = ni, sK = (long)ni * nj;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long nface, base, sD;
    int nd;
    if (D == 0) {
        nface = (long)nj * nk;
        if (t >= nface) return;
        base = sJ * (t % nj) + sK * (t / nj);
        sD = sI;
        nd = ni;
    } else if (D == 1) {
        nface = (long)ni * nk;
        if (t >= nface) return;
        base = (t % ni) + sK * (t / ni);
        sD = sJ;
        nd = nj;
    } else {
        nface = (long)ni * nj;
        if (t >= nface) return;
        base = t;
        sD = sK;
        nd = nk;
    }
    double* q5 = F.q[4] + L.off;
    const int b0 = 2 * D, b1 = 2 * D + 1;
    const long lo = base, hi = base + (long)(nd - 1) * sD;
    switch (C.btype[b0]) {
        case 1: q5[lo] = 0.0; break;
        case 2: case 3: q5[lo] = C.sa[b0][0] * q5[lo + sD] + C.sb[b0][0] * q5[lo + 2 * sD]; break;
        case 6: q5[lo] = q5[base + (long)(nd - 2) * sD]; break;
        default: break;
    }
    switch (C.btype[b1]) {
        case 1: q5[hi] = 0.0; break;
        case 2: case 3: q5[hi] = C.sa[b1][0] * q5[hi - sD] + C.sb[b1][0] * q5[hi - 2 * sD]; break;
        case 6: q5[hi] = q5[base + sD]; break;
        case 8:
            if (D == 0) q5[hi] = q5[hi - sD];
            break;
        default: break;
    }
}

// nu~ floor and xnut = nu~ fv1 on every node (rhs_sa_bcond.F90:279-291)
__global__ void k_sa_clamp(Level L, Fields F, Consts C) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= L.n) return;
    const long g = L.off + t;
    const double q = fmax(1.0e-07, F.q[4][g]);
    F.q[4][g] = q;
    const double chi = q / C.rei;
    const double chic = chi * chi * chi;
    const double fv1 = chic / (chic + cv1c);
    F.xnut[g] = q * fv1;
}

}  // namespace

void launch_des_eddy(hipStream_t s, const Level& L, const Fields& F, const Consts& C, int itk) {
    const int b = 256;
    const unsigned gn = (unsigned)((L.n + b - 1) / b);
    hipLaunchKernelGGL(k_sa_prep, dim3(gn), dim3(b), 0, s, L, F);
    const int bl = 128;
    const unsigned gi = (unsigned)(((long)(L.nj - 2) * (L.nk - 2) + kSL - 1) / kSL);
    const unsigned gj = (unsigned)(((long)(L.ni - 2) * (L.nk - 2) + bl - 1) / bl);
    const unsigned gk = (unsigned)(((long)(L.ni - 2) * (L.nj - 2) + bl - 1) / bl);
    const unsigned fi = (unsigned)(((long)L.nj * L.nk + b - 1) / b);
    const unsigned fj = (unsigned)(((long)L.ni * L.nk + b - 1) / b);
    const unsigned fk = (unsigned)(((long)L.ni * L.nj + b - 1) / b);
    for (int it = 0; it < itk; it++) {
        hipLaunchKernelGGL(k_sa_rhs, dim3(gn), dim3(b), 0, s, L, F, C);
        hipLaunchKernelGGL(k_sa_adi_i, dim3(gi), dim3(kSNT), 0, s, L, F);
        hipLaunchKernelGGL(k_sa_adi<1>, dim3(gj), dim3(bl), 0, s, L, F, C);
        hipLaunchKernelGGL(k_sa_adi<2>, dim3(gk), dim3(bl), 0, s, L, F, C);
        hipLaunchKernelGGL(k_sa_bc<0>, dim3(fi), dim3(b), 0, s, L, F, C);
        hipLaunchKernelGGL(k_sa_bc<1>, dim3(fj), dim3(b), 0, s, L, F, C);
        hipLaunchKernelGGL(k_sa_bc<2>, dim3(fk), dim3(b), 0, s, L, F, C);
        hipLaunchKernelGGL(k_sa_clamp, dim3(gn), dim3(b), 0, s, L, F, C);
    }
}

}  // namespace desalm
