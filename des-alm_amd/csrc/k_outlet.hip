// k_outlet.hip -- characteristic outlet (btype 5) of an i face on the fine
// level: the right-hand side of the boundary plane.
//
// mg_brhs (src/mg_brhs.F90:2-108) assembles rh on the plane from its eta and
// zeta derivatives only: the viscous fluxes of brhs_viscous, the pressure
// dissipation of brhs_diss_p (only its first component; brhs_convec_quick
// zeroes the others), the pressure gradient and divergence of
// brhs_flux_sans_convec (no d/dxi term), QUICK convection (first order on the
// first and last interior rows) of brhs_convec_quick(2), and the BDF2 term
// of brhs_unst_visc_diss.  nonreflect_ibc (src/nonreflect_ibc.F90:1-244) then
// adds the xi-direction terms as characteristic amplitudes L = lambda S^-1
// dq/dxi of the one-sided difference into the domain, with the incoming
// amplitudes set to zero (calc_L_min / calc_L_max by the sign of U).
// One thread per interior (j, k) of the plane; a face flux or dissipation
// term shared by two nodes is evaluated by both, with the same operands.
#include "desalm_internal.h"

namespace desalm {

namespace {

struct Plane {
    Level L;
    int ip;  // 0 or ni - 1
    __device__ long at(int j, int k) const { return L.off + ip + (long)L.ni * (j + (long)L.nj * k); }
};

// viscous flux through the eta face between (j, k) and (j + 1, k)
__device__ void face_eta(const Plane& P, const Fields& F, const double* ret, int j, int k, double de, double dz4,
                         double fv[3]) {
    const long a = P.at(j, k), b = P.at(j + 1, k);
    const long ap = P.at(j, k + 1), am = P.at(j, k - 1), bp = P.at(j + 1, k + 1), bm = P.at(j + 1, k - 1);
    double et[3], ze[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const double* q = F.q[1 + c];
        et[c] = de * (q[b] - q[a]);
        ze[c] = dz4 * (q[bp] - q[bm] + q[ap] - q[am]);
    }
    const double b1 = F.eta[0][b] + F.eta[0][a], b2 = F.eta[1][b] + F.eta[1][a], b3 = F.eta[2][b] + F.eta[2][a];
    const double c1 = F.zet[0][b] + F.zet[0][a], c2 = F.zet[1][b] + F.zet[1][a], c3 = F.zet[2][b] + F.zet[2][a];
    const double uu1 = b1 * et[0] + c1 * ze[0], vv1 = b1 * et[1] + c1 * ze[1], ww1 = b1 * et[2] + c1 * ze[2];
    const double uu2 = b2 * et[0] + c2 * ze[0], vv2 = b2 * et[1] + c2 * ze[1], ww2 = b2 * et[2] + c2 * ze[2];
    const double uu3 = b3 * et[0] + c3 * ze[0], vv3 = b3 * et[1] + c3 * ze[1], ww3 = b3 * et[2] + c3 * ze[2];
    const double c22 = b1 * b1 + b2 * b2 + b3 * b3;
    const double c23 = b1 * c1 + b2 * c2 + b3 * c3;
    const double vtj = 0.25 * (ret[1] + ret[0]) / (F.aj[b] + F.aj[a]);
    fv[0] = vtj * (b1 * uu1 + b2 * vv1 + b3 * ww1 + c22 * et[0] + c23 * ze[0]);
    fv[1] = vtj * (b1 * uu2 + b2 * vv2 + b3 * ww2 + c22 * et[1] + c23 * ze[1]);
    fv[2] = vtj * (b1 * uu3 + b2 * vv3 + b3 * ww3 + c22 * et[2] + c23 * ze[2]);
}

// viscous flux through the zeta face between (j, k) and (j, k + 1)
__device__ void face_zet(const Plane& P, const Fields& F, const double* ret, int j, int k, double de4, double dz,
                         double fv[3]) {
    const long a = P.at(j, k), b = P.at(j, k + 1);
    const long bp = P.at(j + 1, k + 1), bm = P.at(j - 1, k + 1), ap = P.at(j + 1, k), am = P.at(j - 1, k);
    double et[3], ze[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const double* q = F.q[1 + c];
        et[c] = de4 * (q[bp] - q[bm] + q[ap] - q[am]);
        ze[c] = dz * (q[b] - q[a]);
    }
    const double b1 = F.eta[0][b] + F.eta[0][a], b2 = F.eta[1][b] + F.eta[1][a], b3 = F.eta[2][b] + F.eta[2][a];
    const double c1 = F.zet[0][b] + F.zet[0][a], c2 = F.zet[1][b] + F.zet[1][a], c3 = F.zet[2][b] + F.zet[2][a];
    const double uu1 = b1 * et[0] + c1 * ze[0], vv1 = b1 * et[1] + c1 * ze[1], ww1 = b1 * et[2] + c1 * ze[2];
    const double uu2 = b2 * et[0] + c2 * ze[0], vv2 = b2 * et[1] + c2 * ze[1], ww2 = b2 * et[2] + c2 * ze[2];
    const double uu3 = b3 * et[0] + c3 * ze[0], vv3 = b3 * et[1] + c3 * ze[1], ww3 = b3 * et[2] + c3 * ze[2];
    const double c32 = c1 * b1 + c2 * b2 + c3 * b3;
    const double c33 = c1 * c1 + c2 * c2 + c3 * c3;
    const double vtj = 0.25 * (ret[1] + ret[0]) / (F.aj[b] + F.aj[a]);
    fv[0] = vtj * (c1 * uu1 + c2 * vv1 + c3 * ww1 + c32 * et[0] + c33 * ze[0]);
    fv[1] = vtj * (c1 * uu2 + c2 * vv2 + c3 * ww2 + c32 * et[1] + c33 * ze[1]);
    fv[2] = vtj * (c1 * uu3 + c2 * vv3 + c3 * ww3 + c32 * et[2] + c33 * ze[2]);
}

// brhs_diss_p's aux(1, j, k) (d = 0, eta) or aux(2, j, k) (d = 1, zeta) at an
// interior node
__device__ double diss_aux(const Plane& P, const Fields& F, int j, int k, int d, double hsq) {
    const long g = P.at(j, k);
    const double* const* m = d == 0 ? F.eta : F.zet;
    const double gg = m[0][g] * m[0][g] + m[1][g] * m[1][g] + m[2][g] * m[2][g];
    const double tmp1 = gg * F.dtau[g] / F.aj[g];
    const long gp = d == 0 ? P.at(j + 1, k) : P.at(j, k + 1), gm = d == 0 ? P.at(j - 1, k) : P.at(j, k - 1);
    const double tmp2 = hsq * (F.q[0][gp] - 2.0 * F.q[0][g] + F.q[0][gm]);
    return tmp1 * tmp2;
}

__global__ void k_outlet_rhs(Level L, Fields F, Consts C, int inout) {
    const int nj = L.nj, nk = L.nk;
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)(nj - 2) * (nk - 2)) return;
    const int j = 1 + (int)(e % (nj - 2)), k = 1 + (int)(e / (nj - 2));
    const Plane P{L, inout == 1 ? 0 : L.ni - 1};
    const long g = P.at(j, k);
    const double de = L.de, dz = L.dz;
    // brhs_viscous: eta faces j-1/2, j+1/2 and zeta faces k-1/2, k+1/2
    double visc[3];
    {
        double fa[3], fb[3];
        double rt[2];
        rt[0] = C.rei + F.xnut[P.at(j - 1, k)];
        rt[1] = C.rei + F.xnut[g];
        face_eta(P, F, rt, j - 1, k, de, 0.25 * dz, fa);
        rt[0] = rt[1];
        rt[1] = C.rei + F.xnut[P.at(j + 1, k)];
        face_eta(P, F, rt, j, k, de, 0.25 * dz, fb);
#pragma unroll
        for (int c = 0; c < 3; c++) visc[c] = de * (fb[c] - fa[c]);
        rt[0] = C.rei + F.xnut[P.at(j, k - 1)];
        rt[1] = C.rei + F.xnut[g];
        face_zet(P, F, rt, j, k - 1, 0.25 * de, dz, fa);
        rt[0] = rt[1];
        rt[1] = C.rei + F.xnut[P.at(j, k + 1)];
        face_zet(P, F, rt, j, k, 0.25 * de, dz, fb);
#pragma unroll
        for (int c = 0; c < 3; c++) visc[c] = visc[c] + dz * (fb[c] - fa[c]);
    }
    // brhs_diss_p: aux copied from the first / last interior row at the plane's edges
    double diss1;
    {
        const double desq = de * de, dzsq = dz * dz;
        const double a0 = diss_aux(P, F, j, k, 0, desq);
        const double am = j - 1 >= 1 ? diss_aux(P, F, j - 1, k, 0, desq) : a0;
        const double ap = j + 1 <= nj - 2 ? diss_aux(P, F, j + 1, k, 0, desq) : a0;
        const double b0 = diss_aux(P, F, j, k, 1, dzsq);
        const double bm = k - 1 >= 1 ? diss_aux(P, F, j, k - 1, 1, dzsq) : b0;
        const double bp = k + 1 <= nk - 2 ? diss_aux(P, F, j, k + 1, 1, dzsq) : b0;
        diss1 = C.pdiss * ((ap - 2.0 * a0 + am) + (bp - 2.0 * b0 + bm));
    }
    // brhs_flux_sans_convec
    double rh[4];
    {
        const double de2 = 0.5 * de, dz2 = 0.5 * dz;
        const long jp = P.at(j + 1, k), jm = P.at(j - 1, k), kp = P.at(j, k + 1), km = P.at(j, k - 1);
        const double pjp = F.q[0][jp] / F.aj[jp], pjm = F.q[0][jm] / F.aj[jm];
        const double pkp = F.q[0][kp] / F.aj[kp], pkm = F.q[0][km] / F.aj[km];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            rh[1 + c] = de2 * (pjp * F.eta[c][jp] - pjm * F.eta[c][jm]);
            rh[1 + c] = rh[1 + c] + dz2 * (pkp * F.zet[c][kp] - pkm * F.zet[c][km]);
        }
        const long t = g - L.off;  // level-local index of the U/J arrays
        const long tjp = jp - L.off, tjm = jm - L.off, tkp = kp - L.off, tkm = km - L.off;
        (void)t;
        const double dvdy = de2 * (F.ucn[1][tjp] - F.ucn[1][tjm]);
        const double dwdz = dz2 * (F.ucn[2][tkp] - F.ucn[2][tkm]);
        rh[0] = 0.0 + dvdy + dwdz;
    }
    // brhs_convec_quick(2)
#pragma unroll
    for (int d = 0; d < 2; d++) {
        const double h = d == 0 ? de : dz;
        const int p = d == 0 ? j : k, nd = d == 0 ? nj : nk;
        const int dj = d == 0, dk = d == 1;
        const double ucon = 0.5 * F.ucn[1 + d][g - L.off];
        const double up = ucon + fabs(ucon), um = ucon - fabs(ucon);
        const long g1p = P.at(j + dj, k + dk), g1m = P.at(j - dj, k - dk);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const double* q = F.q[1 + c];
            double r;
            if (p >= 2 && p <= nd - 3) {
                const long g2p = P.at(j + 2 * dj, k + 2 * dk), g2m = P.at(j - 2 * dj, k - 2 * dk);
                r = 0.125 * h *
                    (up * (3.0 * q[g1p] + 3.0 * q[g] - 7.0 * q[g1m] + q[g2m]) +
                     um * (-3.0 * q[g1m] - 3.0 * q[g] + 7.0 * q[g1p] - q[g2p]));
            } else {
                r = h * (up * (q[g] - q[g1m]) + um * (q[g1p] - q[g]));
            }
            rh[1 + c] = rh[1 + c] + r;
        }
    }
    // brhs_unst_visc_diss
    {
        const double dtaj = 2.0 * C.delti * F.aj[g];
        rh[0] = rh[0] + diss1;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const double dqdt = C.e_source * (3.0 * F.q[1 + c][g] - 4.0 * F.qn[1 + c][g] + F.qnm1[1 + c][g]) / dtaj;
            rh[1 + c] = rh[1 + c] + 0.0 + dqdt - visc[c];
        }
    }
    // nonreflect_ibc
    {
        const long nb = g + (inout == 1 ? 1 : -1);
        const double u = F.q[1][g], v = F.q[2][g], w = F.q[3][g];
        const double aj = F.aj[g];
        const double ucon = F.ucn[0][g - L.off] * aj;
        const double c1 = F.csi[0][g], c2 = F.csi[1][g], c3 = F.csi[2][g];
        const double gjj = c1 * c1 + c2 * c2 + c3 * c3;
        const double sg = sqrt(gjj);
        const double cj = sqrt(ucon * ucon + gjj);
        double dq[4];
#pragma unroll
        for (int m = 0; m < 4; m++) dq[m] = inout == 1 ? F.q[m][nb] - F.q[m][g] : F.q[m][g] - F.q[m][nb];
        const double lam1 = ucon - cj, lam2 = ucon + cj, lam3 = ucon, lam4 = ucon;
        const double kj = c1 + c2 + c3;
        const double d1 = (c2 - c1 + lam1 * (v - u)) / kj;
        const double d2 = (c2 - c1 + lam2 * (v - u)) / kj;
        const double d3 = (c3 - c1 + lam1 * (w - u)) / kj;
        const double d4 = (c3 - c1 + lam2 * (w - u)) / kj;
        const double d5 = (c2 - c1 + lam3 * (v - u)) / kj * 2.0;
        const double d6 = (c3 - c1 + lam4 * (w - u)) / kj * 2.0;
        const double t2 = 2.0 * cj * cj / kj;
        const double si[4][4] = {{-lam2, c1, c2, c3},
                                 {-lam1, c1, c2, c3},
                                 {lam2 * d1 + lam1 * d2, -t2 - c1 * d5, t2 - c2 * d5, -c3 * d5},
                                 {lam2 * d3 + lam1 * d4, -t2 - c1 * d6, -c2 * d6, t2 - c3 * d6}};
        const double den = 2.0 * cj * cj;
        double lg[4];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            double acc = 0.0;
#pragma unroll
            for (int c = 0; c < 4; c++) acc = acc + si[r][c] * sg / den * dq[c];
            lg[r] = acc;
        }
        lg[0] = lam1 * lg[0];
        lg[1] = lam2 * lg[1];
        lg[2] = lam3 * lg[2];
        lg[3] = lam4 * lg[3];
        double L1, L2, L3, L4;
        if (inout == 1) {
            L1 = lg[0];
            L2 = 0.0;
            L3 = ucon < 0.0 ? lg[2] : 0.0;
            L4 = ucon < 0.0 ? lg[3] : 0.0;
        } else {
            L1 = 0.0;
            L2 = lg[1];
            L3 = ucon < 0.0 ? 0.0 : lg[2];
            L4 = ucon < 0.0 ? 0.0 : lg[3];
        }
        const double r1 = (-cj * L1 + cj * L2) / sg;
        const double r2 = ((c1 + u * lam1) * L1 + (c1 + u * lam2) * L2 - c2 * L3 - c3 * L4) / sg;
        const double r3 = ((c2 + v * lam1) * L1 + (c2 + v * lam2) * L2 + (c1 + c3) * L3 - c3 * L4) / sg;
        const double r4 = ((c3 + w * lam1) * L1 + (c3 + w * lam2) * L2 - c2 * L3 + (c1 + c2) * L4) / sg;
        rh[0] = rh[0] + r1 / aj;
        rh[1] = rh[1] + r2 / aj;
        rh[2] = rh[2] + r3 / aj;
        rh[3] = rh[3] + r4 / aj;
    }
#pragma unroll
    for (int m = 0; m < 4; m++) F.rh[m][g] = rh[m];
}

}  // namespace

void launch_outlet_rhs(hipStream_t s, const Level& L, const Fields& F, const Consts& C) {
    const long n = (long)(L.nj - 2) * (L.nk - 2);
    const unsigned nb = (unsigned)((n + 127) / 128);
    if (C.btype[0] == DESALM_BC_OUTLET_MOC) hipLaunchKernelGGL(k_outlet_rhs, dim3(nb), dim3(128), 0, s, L, F, C, 1);
    if (C.btype[1] == DESALM_BC_OUTLET_MOC) hipLaunchKernelGGL(k_outlet_rhs, dim3(nb), dim3(128), 0, s, L, F, C, 2);
}

}  // namespace desalm
