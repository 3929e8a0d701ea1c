// k_residual.hip -- residual assembly of one DAF smoother pass on one level.
//
// Three kernels produce the right-hand side of solver_daf steps 2-10
// (src/solver_daf.F90:199-315):
//   k_prep     dtau and the pressure dissipation's fv per node
//   k_residual central pressure gradient and divergence, QUICK or first-order
//              convection, 4th-difference pressure dissipation, BDF2 term,
//              zeroing of the boundary planes
//   k_visc_pk  2nd-order viscous fluxes and the FAS forcing
// No face arrays touch HBM: k_residual's face terms are evaluated by both
// nodes that share them (one thread per node gathers its stencil from the
// SoA fields); k_visc_pk stages its tiles in LDS and shares its face fluxes
// there.
// Operand order inside every expression follows the reference routine cited.
#include <algorithm>

#include "desalm_internal.h"

namespace desalm {

namespace {

// Per-node inputs of the residual, every node of the level: dtau of
// rhs_daf_dtau (src/rhs_daf_dtau.F90:51-83), or of met_daf_dtau
// (src/met_daf_dtau.F90:53-93) when C.dtau_met, and the fv of rhs_diss_p per
// direction (rhs_diss_p.F90:58-117) at the nodes interior in that direction
// into F.fv (level-local): the residual reads it at a node and its two
// neighbours.  For the characteristic outlet (k_outlet.hip), which reads
// them, also U/J (rhs_contra_j.F90:14-29) into F.ucn (mode & kPrepUcn; not
// where the SA update has just stored U/J of the same velocities) and p/J
// (rhs_flux_sans_convec.F90:60) into F.pj (mode & kPrepPj), level-local; the
// residual forms both where it uses them.
constexpr int kPrepUcn = 1, kPrepPj = 2;
__host__ __device__ inline int prep_mode(const Consts& C, int n, bool ucn_kept) {
    const bool outlet = n == 1 && (C.btype[0] == DESALM_BC_OUTLET_MOC || C.btype[1] == DESALM_BC_OUTLET_MOC);
    return outlet ? kPrepPj | (ucn_kept ? 0 : kPrepUcn) : 0;
}
__device__ __forceinline__ void prep_body(const Level& L, const Fields& F, const Consts& C, int mode, unsigned bid,
                                          unsigned nblk) {
    const bool with_ucn = mode & kPrepUcn, with_pj = mode & kPrepPj;
    const long t = (long)xcd_block(bid, nblk) * blockDim.x + threadIdx.x;
    if (t >= L.n) return;
    const long g = L.off + t;
    const int ni = L.ni, nj = L.nj;
    const int i = (int)(t % ni), j = (int)((t / ni) % nj), k = (int)(t / ((long)ni * nj));
    const int pos[3] = {i, j, k}, nd[3] = {ni, nj, L.nk};
    const long st[3] = {1, (long)ni, (long)ni * nj};
    // every load is issued before the first store; the pressure neighbours
    // without a branch (a node on a boundary plane of direction D reads
    // itself there and stores no fv of D)
    const double c1 = F.csi[0][g], c2 = F.csi[1][g], c3 = F.csi[2][g];
    const double e1 = F.eta[0][g], e2 = F.eta[1][g], e3 = F.eta[2][g];
    const double z1 = F.zet[0][g], z2 = F.zet[1][g], z3 = F.zet[2][g];
    const double aj = F.aj[g];
    double u = 0.0, v = 0.0, w = 0.0;
    if (with_ucn) {
        u = F.q[1][g];
        v = F.q[2][g];
        w = F.q[3][g];
    }
    const double* p = F.q[0];
    const double p0 = p[g];
    double pm[3], pp[3];
#pragma unroll
    for (int D = 0; D < 3; D++) {
        pm[D] = p[g - (pos[D] >= 1 ? st[D] : 0)];
        pp[D] = p[g + (pos[D] <= nd[D] - 2 ? st[D] : 0)];
    }
    const double xnut = C.dtau_met ? 0.0 : F.xnut[g];
    if (with_ucn) {
        const double tmp = 1.0 / aj;
        F.ucn[0][t] = tmp * (c1 * u + c2 * v + c3 * w);
        F.ucn[1][t] = tmp * (e1 * u + e2 * v + e3 * w);
        F.ucn[2][t] = tmp * (z1 * u + z2 * v + z3 * w);
    }
    if (with_pj) F.pj[t] = p0 / aj;
    const double g1 = c1 * c1 + c2 * c2 + c3 * c3;
    const double g2 = e1 * e1 + e2 * e2 + e3 * e3;
    const double g3 = z1 * z1 + z2 * z2 + z3 * z3;
    const double dti1 = C.cfl1 / sqrt(g1) / L.dc;
    const double dti2 = C.cfl1 / sqrt(g2) / L.de;
    const double dti3 = C.cfl1 / sqrt(g3) / L.dz;
    double dtv1, dtv2, dtv3;
    if (C.dtau_met) {
        const double vr = C.vnn1 * C.ren;
        dtv1 = vr / (L.dc * L.dc) / g1;
        dtv2 = vr / (L.de * L.de) / g2;
        dtv3 = vr / (L.dz * L.dz) / g3;
    } else {
        const double ret = C.rei + xnut;
        dtv1 = C.vnn1 / ret / (L.dc * L.dc) / g1;
        dtv2 = C.vnn1 / ret / (L.de * L.de) / g2;
        dtv3 = C.vnn1 / ret / (L.dz * L.dz) / g3;
    }
    double d = fmin(dti1, dti2);
    d = fmin(d, dti3);
    d = fmin(d, dtv1);
    d = fmin(d, dtv2);
    d = fmin(d, dtv3);
    F.dtau[g] = d;
    // fv = |grad xi_D|^2 dtau / J * h^2 (p(+1) - 2 p + p(-1))
    const double gg[3] = {g1, g2, g3};
    const double hh[3] = {L.dc, L.de, L.dz};
#pragma unroll
    for (int D = 0; D < 3; D++) {
        if (pos[D] < 1 || pos[D] > nd[D] - 2) continue;
        const double tmp1 = gg[D] * d / aj;
        const double tmp2 = (hh[D] * hh[D]) * (pp[D] - 2.0 * p0 + pm[D]);
        F.fv[D][t] = tmp1 * tmp2;
    }
}

// QUICK (rhs_convec_quick.F90:89-201 and its eta/zet copies) or first-order
// upwind term of one velocity component along one direction, for rh(2:4):
// ucon is half of U/J of that direction at the node, h the grid spacing,
// um2 .. up2 the component at offsets -2 .. +2 along the line (where
// first_order, um2 and up2 hold the -1 and +1 values and are unused).  Both
// forms are evaluated and one is selected, so that every load is issued up
// front.
__device__ __forceinline__ double convec_term(double ucon, double h, bool first_order, double um2, double um1,
                                              double u0, double up1, double up2) {
    const double up = ucon + fabs(ucon), um = ucon - fabs(ucon);
    const double fo = h * (up * (u0 - um1) + um * (up1 - u0));
    const double qk = 0.125 * h *
                      (up * (3.0 * up1 + 3.0 * u0 - 7.0 * um1 + um2) +
                       um * (-3.0 * um1 - 3.0 * u0 + 7.0 * up1 - up2));
    return first_order ? fo : qk;
}

// U/J of the direction whose metric row is m (rhs_contra_j.F90:14-29)
__device__ __forceinline__ double contra_j(double aj, const double (&m)[3], const double (&u)[3]) {
    const double tmp = 1.0 / aj;
    return tmp * (m[0] * u[0] + m[1] * u[1] + m[2] * u[2]);
}

// Residual of solver_daf on one level without its viscous term and FAS
// forcing (k_visc_pk adds those).  recalc = 0: first evaluation of the pass
// (diss(1) fresh and stored); recalc = 1: the decide_recalc_rh branch
// (src/solver_daf.F90:432-484), which reuses the stored diss(1) of the
// pre-update q and dtau.  p/J (rhs_flux_sans_convec.F90:60) at the six
// neighbours and U/J (rhs_contra_j.F90:14-29) at the node and its six
// neighbours are formed from p, J, the metrics and the velocities, which the
// convection and the pressure gradient gather anyway.
__device__ __forceinline__ void residual_body(const Level& L, const Fields& F, const Consts& C, int n, int recalc,
                                              unsigned bid, unsigned nblk) {
    const long t = (long)xcd_block(bid, nblk) * blockDim.x + threadIdx.x;
    if (t >= L.n) return;
    const int ni = L.ni, nj = L.nj, nk = L.nk;
    const int i = (int)(t % ni);
    const int j = (int)((t / ni) % nj);
    const int k = (int)(t / ((long)ni * nj));
    const long g = L.off + t;
    if (i == 0 || i == ni - 1 || j == 0 || j == nj - 1 || k == 0 || k == nk - 1) {
        // rh = 0 on the boundary planes; pk there is left as it is
        // (rhs_pk adds rh = 0 to it, then the plane is zeroed, :301-306)
#pragma unroll
        for (int m = 0; m < 4; m++) F.rh[m][g] = 0.0;
        return;
    }
    const long sD[3] = {1, (long)ni, (long)ni * nj};
    const int pos[3] = {i, j, k}, len[3] = {ni, nj, nk};
    const double dc2 = 0.5 * L.dc, de2 = 0.5 * L.de, dz2 = 0.5 * L.dz;
    const double* const* met[3] = {F.csi, F.eta, F.zet};

    // every load first.  Neighbour q = 2 D + s (s = 0: -1, s = 1: +1 along
    // D): p, J, the metric row of D and the velocities
    double pn[6], ajn[6], mn[6][3], un[6][3];
#pragma unroll
    for (int D = 0; D < 3; D++)
#pragma unroll
        for (int s = 0; s < 2; s++) {
            const int q = 2 * D + s;
            const long gn = g + (s ? sD[D] : -sD[D]);
            pn[q] = F.q[0][gn];
            ajn[q] = F.aj[gn];
#pragma unroll
            for (int c = 0; c < 3; c++) mn[q][c] = met[D][c][gn];
#pragma unroll
            for (int m = 0; m < 3; m++) un[q][m] = F.q[m + 1][gn];
        }
    // the node
    const double aj0 = F.aj[g];
    double m0[3][3], u0[3];
#pragma unroll
    for (int D = 0; D < 3; D++)
#pragma unroll
        for (int c = 0; c < 3; c++) m0[D][c] = met[D][c][g];
#pragma unroll
    for (int m = 0; m < 3; m++) u0[m] = F.q[m + 1][g];
    // the velocities two nodes away along each line for QUICK: on the fine
    // level but at the first and last interior node (where the +-2 loads
    // read the +-1 neighbours; the +-2 ones may lie outside the level)
    const bool quick = n == 1;
    bool fo[3];
    double u2[3][2][3];
#pragma unroll
    for (int D = 0; D < 3; D++) {
        fo[D] = !quick || pos[D] == 1 || pos[D] == len[D] - 2;
        const long s2 = fo[D] ? sD[D] : 2 * sD[D];
#pragma unroll
        for (int m = 0; m < 3; m++) {
            u2[D][0][m] = F.q[m + 1][g - s2];
            u2[D][1][m] = F.q[m + 1][g + s2];
        }
    }
    double fvn[9], d1 = 0.0;
    if (!recalc) {
        // the boundary copies of fv (rhs_diss_p.F90:124-129) replace the
        // out-of-domain neighbour by the node itself
#pragma unroll
        for (int D = 0; D < 3; D++) {
            const double* fv = F.fv[D];
            fvn[3 * D] = fv[t - (pos[D] == 1 ? 0 : sD[D])];
            fvn[3 * D + 1] = fv[t];
            fvn[3 * D + 2] = fv[t + (pos[D] == len[D] - 2 ? 0 : sD[D])];
        }
    } else {
#pragma unroll
        for (int c = 0; c < 9; c++) fvn[c] = 0.0;
        d1 = F.diss1[g];
    }
    double qnv[3], qmv[3];
#pragma unroll
    for (int m = 0; m < 3; m++) {
        qnv[m] = F.qn[m + 1][g];
        qmv[m] = F.qnm1[m + 1][g];
    }

    // p/J and U/J of the neighbours
    double pjn[6], ucn[6];
#pragma unroll
    for (int q = 0; q < 6; q++) {
        pjn[q] = pn[q] / ajn[q];
        ucn[q] = contra_j(ajn[q], mn[q], un[q]);
    }
    // central pressure gradient (rhs_flux_sans_convec.F90:75-116)
    double r[4];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        double v = dc2 * (pjn[1] * mn[1][c] - pjn[0] * mn[0][c]);
        v = v + de2 * (pjn[3] * mn[3][c] - pjn[2] * mn[2][c]);
        v = v + dz2 * (pjn[5] * mn[5][c] - pjn[4] * mn[4][c]);
        r[c + 1] = v;
    }
    // divergence of U/J (:120-135)
    {
        const double dudx = dc2 * (ucn[1] - ucn[0]);
        const double dvdy = de2 * (ucn[3] - ucn[2]);
        const double dwdz = dz2 * (ucn[5] - ucn[4]);
        r[0] = dudx + dvdy + dwdz;
    }
    // convection: QUICK on the fine level, first-order upwind on coarse levels
#pragma unroll
    for (int D = 0; D < 3; D++) {
        const double h = D == 0 ? L.dc : D == 1 ? L.de : L.dz;
        const double ucon = 0.5 * contra_j(aj0, m0[D], u0);
#pragma unroll
        for (int m = 0; m < 3; m++)
            r[m + 1] = r[m + 1] + convec_term(ucon, h, fo[D], u2[D][0][m], un[2 * D][m], u0[m], un[2 * D + 1][m],
                                              u2[D][1][m]);
    }
    // pressure dissipation (rhs_diss_p.F90:58-203); the boundary copies of
    // fv (:124-129) replace the out-of-domain neighbour by the node itself
    double diss1;
    if (!recalc) {
        diss1 = C.pdiss * ((fvn[2] - 2.0 * fvn[1] + fvn[0]) + (fvn[5] - 2.0 * fvn[4] + fvn[3]) +
                           (fvn[8] - 2.0 * fvn[7] + fvn[6]));
        F.diss1[g] = diss1;
    } else {
        diss1 = d1;
    }
    // unsteady and dissipation terms (rhs_unst_visc_diss.F90:29-48);
    // diss(2:4) is zero (rhs_convec_quick.F90:433); then the body force
    // (no counterpart in the reference), -f / J
    {
        const double dtaj = 2.0 * C.delti * aj0;
        r[0] = r[0] + diss1;
#pragma unroll
        for (int m = 1; m < 4; m++) {
            const double dqdt = C.e_source * (3.0 * u0[m - 1] - 4.0 * qnv[m - 1] + qmv[m - 1]) / dtaj;
            r[m] = r[m] + 0.0 + dqdt;
        }
        if (C.fb[0] != 0.0 || C.fb[1] != 0.0 || C.fb[2] != 0.0)
#pragma unroll
            for (int m = 1; m < 4; m++) r[m] = r[m] - C.fb[m - 1] / aj0;
    }
#pragma unroll
    for (int m = 0; m < 4; m++) F.rh[m][g] = r[m];
}

// Viscous term subtracted from rh(2:4) (rhs_viscous.F90:112-329,
// rhs_unst_visc_diss.F90:45-47) and the FAS forcing of coarse levels
// (rhs_pk.F90:35-65): pk = pk - rh when calc_pk, then rh = rh + pk.
//
// A workgroup of kTX x kTY threads owns an (i, j) tile of (kTX-1) x (kTY-1)
// interior nodes plus one halo column and row on the low side, and marches
// through kc consecutive k planes.  Every thread evaluates the +i face of its
// node (rows with an output) and the +j face (columns with an output) once
// and shares them through LDS.  u, v, w of three planes (k-1, k, k+1) and
// the metrics, J and nu_t of plane k sit in LDS for the tile plus one high
// halo column and row (kLX x kLY nodes), so every face reads its 19-point
// velocity footprint and both nodes' metrics from LDS instead of issuing
// global loads.  The k face above a node is carried into the next plane (each
// k face is evaluated once).  The next planes are loaded into registers while
// the current one is computed.
constexpr int kTX = 32, kTY = 8;
constexpr int kLX = kTX + 1, kLY = kTY + 1, kLN = kLX * kLY;
constexpr int kHalo = kLN - kTX * kTY;  // 41 halo nodes: column kTX, then row kTY
constexpr int kMet = 11;                // csi 3, eta 3, zet 3, aj, xnut

struct Met {
    double v[kMet];
};

__device__ __forceinline__ void load_met(const Fields& F, long g, bool ok, Met& m) {
#pragma unroll
    for (int c = 0; c < 3; c++) {
        m.v[c] = ok ? F.csi[c][g] : 0.0;
        m.v[3 + c] = ok ? F.eta[c][g] : 0.0;
        m.v[6 + c] = ok ? F.zet[c][g] : 0.0;
    }
    m.v[9] = ok ? F.aj[g] : 0.0;
    m.v[10] = ok ? F.xnut[g] : 0.0;
}

// One viscous face flux (rhs_viscous.F90:112-163 for D=0, :189-243 for D=1,
// :263-316 for D=2) between node 0 and node 1 = node 0 + e_D; U(m, a, b, c)
// returns velocity component m at node 0 + (a, b, c), m0 / m1 are the two
// nodes' metrics, J and nu_t.  uu[r][m] is the r-th metric row applied to the
// gradient of velocity component m; flux component m is the face metric
// applied to uu[.][m] plus cc . grad(u_m) (rhs_viscous.F90:158-163).
template <int D, typename U>
__device__ __forceinline__ void visc_face_t(const Level& L, const Consts& C, U u, const Met& m0, const Met& m1,
                                            double fv[3]) {
    constexpr int X = D == 0, Y = D == 1, Z = D == 2;
    const double hc = L.dc, he = L.de, hz = L.dz;
    const double hc4 = 0.25 * hc, he4 = 0.25 * he, hz4 = 0.25 * hz;
    double dcs[3], det[3], dze[3];
#pragma unroll
    for (int m = 0; m < 3; m++) {
        if (D == 0) {
            dcs[m] = hc * (u(m, 1, 0, 0) - u(m, 0, 0, 0));
            det[m] = he4 * (u(m, 1, 1, 0) - u(m, 1, -1, 0) + u(m, 0, 1, 0) - u(m, 0, -1, 0));
            dze[m] = hz4 * (u(m, 1, 0, 1) - u(m, 1, 0, -1) + u(m, 0, 0, 1) - u(m, 0, 0, -1));
        } else if (D == 1) {
            dcs[m] = hc4 * (u(m, 1, 1, 0) - u(m, -1, 1, 0) + u(m, 1, 0, 0) - u(m, -1, 0, 0));
            det[m] = he * (u(m, 0, 1, 0) - u(m, 0, 0, 0));
            dze[m] = hz4 * (u(m, 0, 1, 1) - u(m, 0, 1, -1) + u(m, 0, 0, 1) - u(m, 0, 0, -1));
        } else {
            dcs[m] = hc4 * (u(m, 1, 0, 1) - u(m, -1, 0, 1) + u(m, 1, 0, 0) - u(m, -1, 0, 0));
            det[m] = he4 * (u(m, 0, 1, 1) - u(m, 0, -1, 1) + u(m, 0, 1, 0) - u(m, 0, -1, 0));
            dze[m] = hz * (u(m, 0, 0, 1) - u(m, 0, 0, 0));
        }
    }
    (void)X; (void)Y; (void)Z;
    const double a1 = m1.v[0] + m0.v[0], a2 = m1.v[1] + m0.v[1], a3 = m1.v[2] + m0.v[2];
    const double b1 = m1.v[3] + m0.v[3], b2 = m1.v[4] + m0.v[4], b3 = m1.v[5] + m0.v[5];
    const double c1 = m1.v[6] + m0.v[6], c2 = m1.v[7] + m0.v[7], c3 = m1.v[8] + m0.v[8];
    double uu[3][3];
#pragma unroll
    for (int m = 0; m < 3; m++) {
        uu[0][m] = a1 * dcs[m] + b1 * det[m] + c1 * dze[m];
        uu[1][m] = a2 * dcs[m] + b2 * det[m] + c2 * dze[m];
        uu[2][m] = a3 * dcs[m] + b3 * det[m] + c3 * dze[m];
    }
    const double f1 = D == 0 ? a1 : D == 1 ? b1 : c1;
    const double f2 = D == 0 ? a2 : D == 1 ? b2 : c2;
    const double f3 = D == 0 ? a3 : D == 1 ? b3 : c3;
    const double cca = f1 * a1 + f2 * a2 + f3 * a3;
    const double ccb = f1 * b1 + f2 * b2 + f3 * b3;
    const double ccc = f1 * c1 + f2 * c2 + f3 * c3;
    const double ret1 = C.rei + m1.v[10], ret0 = C.rei + m0.v[10];
    const double vtj = 0.25 * (ret1 + ret0) / (m1.v[9] + m0.v[9]);
#pragma unroll
    for (int r = 0; r < 3; r++)
        fv[r] = vtj * (f1 * uu[r][0] + f2 * uu[r][1] + f3 * uu[r][2] + cca * dcs[r] + ccb * det[r] + ccc * dze[r]);
}

// tiles of the viscous kernel along i and j
__host__ __device__ inline unsigned visc_gx(const Level& L) { return (unsigned)((L.ni - 2 + kTX - 2) / (kTX - 1)); }
__host__ __device__ inline unsigned visc_gy(const Level& L) { return (unsigned)((L.nj - 2 + kTY - 2) / (kTY - 1)); }

// bid: linear workgroup index over (i tile, j tile, k chunk), i fastest
__device__ __forceinline__ void visc_pk_body(const Level& L, const Fields& F, const Consts& C, int n, int calc_pk,
                                             int kc, unsigned bid, unsigned nblk) {
    __shared__ double s_u[3][3][kLN];  // velocity component, plane slot (k mod 3), tile node
    __shared__ double s_m[kMet][kLN];  // metrics, J, nu_t of the current plane
    __shared__ double s_fi[3][kTY][kTX];
    __shared__ double s_fj[3][kTY][kTX];
    const int ni = L.ni, nj = L.nj, nk = L.nk;
    const unsigned gx = visc_gx(L), gy = visc_gy(L);
    const unsigned lb = xcd_block(bid, nblk);
    const int bx = (int)(lb % gx), by = (int)((lb / gx) % gy), bz = (int)(lb / (gx * gy));
    const int tid = (int)threadIdx.x;
    const int tx = tid % kTX, ty = tid / kTX;
    const int i0 = bx * (kTX - 1), j0 = by * (kTY - 1);
    const int kb = 1 + bz * kc, ke = min(kb + kc, nk - 1);  // output planes kb .. ke-1
    const long sJ = ni, sK = (long)ni * nj;
    // the thread's own node, and (tid < kHalo) one halo node of the tile
    const int own = ty * kLX + tx;
    const int hr = tid < kLY ? tid : kTY, hc = tid < kLY ? kTX : tid - kLY;
    const int halo = hr * kLX + hc;
    const bool has_halo = tid < kHalo;
    const int i = i0 + tx, j = j0 + ty;
    const bool own_ok = i < ni && j < nj;
    const bool halo_ok = has_halo && i0 + hc < ni && j0 + hr < nj;
    const long col = L.off + (long)i + sJ * j;
    const long hcol = L.off + (long)(i0 + hc) + sJ * (j0 + hr);
    const bool in_i = i <= ni - 2, in_j = j <= nj - 2;
    const bool out = tx >= 1 && ty >= 1 && in_i && in_j;
    const bool face_i = ty >= 1 && in_i && in_j;
    const bool face_j = tx >= 1 && in_i && in_j;

    auto load_uvw = [&](int k, double ro[3], double rh_[3]) {
#pragma unroll
        for (int m = 0; m < 3; m++) {
            ro[m] = own_ok ? F.q[m + 1][col + sK * k] : 0.0;
            rh_[m] = halo_ok ? F.q[m + 1][hcol + sK * k] : 0.0;
        }
    };
    auto store_uvw = [&](int k, const double ro[3], const double rh_[3]) {
        const int sl = k % 3;
#pragma unroll
        for (int m = 0; m < 3; m++) {
            s_u[m][sl][own] = ro[m];
            if (has_halo) s_u[m][sl][halo] = rh_[m];
        }
    };
    auto store_met = [&](const Met& mo, const Met& mh) {
#pragma unroll
        for (int c = 0; c < kMet; c++) {
            s_m[c][own] = mo.v[c];
            if (has_halo) s_m[c][halo] = mh.v[c];
        }
    };
    auto met_at = [&](int node) {
        Met m;
#pragma unroll
        for (int c = 0; c < kMet; c++) m.v[c] = s_m[c][node];
        return m;
    };

    // prologue: planes kb-1, kb, kb+1 of u, v, w; metrics of plane kb
    {
        double ro[3], rh_[3];
        for (int k = kb - 1; k <= kb + 1; k++) {
            load_uvw(k, ro, rh_);
            store_uvw(k, ro, rh_);
        }
    }
    Met mcur, mh;
    load_met(F, col + sK * kb, own_ok, mcur);
    load_met(F, hcol + sK * kb, halo_ok, mh);
    store_met(mcur, mh);
    double fk[3] = {0.0, 0.0, 0.0};
    {
        Met mlow;
        load_met(F, col + sK * (kb - 1), own_ok, mlow);
        __syncthreads();
        if (out) {
            const int sl0 = (kb - 1) % 3, sl1 = kb % 3;
            auto u = [&](int m, int a, int b, int c) {
                return s_u[m][c == 0 ? sl0 : sl1][own + b * kLX + a];
            };
            visc_face_t<2>(L, C, u, mlow, mcur, fk);
        }
    }
    for (int k = kb; k < ke; k++) {
        // next planes into registers: metrics of k+1, velocities of k+2
        Met mnext, mhn;
        load_met(F, col + sK * (k + 1), own_ok, mnext);
        load_met(F, hcol + sK * (k + 1), halo_ok, mhn);
        double ro[3] = {0.0, 0.0, 0.0}, rh_[3] = {0.0, 0.0, 0.0};
        const bool more = k + 1 < ke;
        if (more) load_uvw(k + 2, ro, rh_);
        // this plane's rh as well, so that its loads overlap the face fluxes
        const long g = col + sK * k;
        double r[4] = {0.0, 0.0, 0.0, 0.0};
        if (out)
#pragma unroll
            for (int m = 0; m < 4; m++) r[m] = F.rh[m][g];
        const int slm = (k + 2) % 3, sl0 = k % 3, slp = (k + 1) % 3;
        auto uk = [&](int m, int a, int b, int c) {
            return s_u[m][c < 0 ? slm : c == 0 ? sl0 : slp][own + b * kLX + a];
        };
        if (face_i) {
            double f[3];
            visc_face_t<0>(L, C, uk, mcur, met_at(own + 1), f);
#pragma unroll
            for (int c = 0; c < 3; c++) s_fi[c][ty][tx] = f[c];
        }
        if (face_j) {
            double f[3];
            visc_face_t<1>(L, C, uk, mcur, met_at(own + kLX), f);
#pragma unroll
            for (int c = 0; c < 3; c++) s_fj[c][ty][tx] = f[c];
        }
        double fkp[3] = {0.0, 0.0, 0.0};
        if (out) visc_face_t<2>(L, C, uk, mcur, mnext, fkp);
        __syncthreads();
        if (out) {
            double visc[3];
#pragma unroll
            for (int c = 0; c < 3; c++) visc[c] = L.dc * (s_fi[c][ty][tx] - s_fi[c][ty][tx - 1]);
#pragma unroll
            for (int c = 0; c < 3; c++) visc[c] = visc[c] + L.de * (s_fj[c][ty][tx] - s_fj[c][ty - 1][tx]);
#pragma unroll
            for (int c = 0; c < 3; c++) visc[c] = visc[c] + L.dz * (fkp[c] - fk[c]);
#pragma unroll
            for (int m = 1; m < 4; m++) r[m] = r[m] - visc[m - 1];
            if (n > 1) {
                double pk[4];  // all four loaded before the first store
#pragma unroll
                for (int m = 0; m < 4; m++) pk[m] = F.pk[m][g];
#pragma unroll
                for (int m = 0; m < 4; m++) {
                    if (calc_pk == 1) {
                        pk[m] = pk[m] - r[m];
                        F.pk[m][g] = pk[m];
                    }
                    r[m] = r[m] + pk[m];
                }
            }
#pragma unroll
            for (int m = 0; m < 4; m++) F.rh[m][g] = r[m];
        }
        // plane k+1 becomes current: its metrics into LDS, plane k+2's
        // velocities into the slot of plane k-1 (no barrier needed before:
        // the part above reads only the face fluxes)
        store_met(mnext, mhn);
        if (more) store_uvw(k + 2, ro, rh_);
        mcur = mnext;
#pragma unroll
        for (int c = 0; c < 3; c++) fk[c] = fkp[c];
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256) k_prep(Level L, Fields F, Consts C, int mode) {
    prep_body(L, F, C, mode, blockIdx.x, gridDim.x);
}
__global__ void __launch_bounds__(256) k_prep_z(ZoneGrid G, int n, int ucn_kept) {
    const ZoneBlock b = zone_block(G);
    const ZoneDesc& Z = G.z[b.z];
    prep_body(Z.lv[n], Z.F, Z.C, prep_mode(Z.C, n, ucn_kept), b.bid, b.nblk);
}
__global__ void __launch_bounds__(256) k_residual(Level L, Fields F, Consts C, int n, int recalc) {
    residual_body(L, F, C, n, recalc, blockIdx.x, gridDim.x);
}
__global__ void __launch_bounds__(256) k_residual_z(ZoneGrid G, int n, int recalc) {
    const ZoneBlock b = zone_block(G);
    const ZoneDesc& Z = G.z[b.z];
    residual_body(Z.lv[n], Z.F, Z.C, n, recalc, b.bid, b.nblk);
}
__global__ void __launch_bounds__(kTX * kTY, 2) k_visc_pk(Level L, Fields F, Consts C, int n, int calc_pk, int kc) {
    visc_pk_body(L, F, C, n, calc_pk, kc, blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z),
                 gridDim.x * gridDim.y * gridDim.z);
}
__global__ void __launch_bounds__(kTX * kTY, 2) k_visc_pk_z(ZoneGrid G, int n, int calc_pk, int kc) {
    const ZoneBlock b = zone_block(G);
    const ZoneDesc& Z = G.z[b.z];
    visc_pk_body(Z.lv[n], Z.F, Z.C, n, calc_pk, kc, b.bid, b.nblk);
}

}  // namespace

void launch_dtau(hipStream_t s, const Level& L, const Fields& F, const Consts& C, int n, bool ucn_kept) {
    const int b = 256;
    hipLaunchKernelGGL(k_prep, dim3((unsigned)((L.n + b - 1) / b)), dim3(b), 0, s, L, F, C,
                       prep_mode(C, n, ucn_kept));
}

static unsigned node_blocks(const Level& L) { return (unsigned)((L.n + 255) / 256); }

void launch_dtau_z(hipStream_t s, const ZoneList& Z, int n, bool ucn_kept) {
    unsigned nb = 0;
    const ZoneGrid G = Z.grid([&](const ZoneDesc& z) { return node_blocks(z.lv[n]); }, nb);
    hipLaunchKernelGGL(k_prep_z, dim3(nb), dim3(256), 0, s, G, n, (int)ucn_kept);
}

void launch_residual_z(hipStream_t s, const ZoneList& Z, int n, int recalc, int calc_pk, int kc) {
    unsigned nb = 0;
    const ZoneGrid G = Z.grid([&](const ZoneDesc& z) { return node_blocks(z.lv[n]); }, nb);
    hipLaunchKernelGGL(k_residual_z, dim3(nb), dim3(256), 0, s, G, n, recalc);
    // one plane count for every zone: the smallest zone's default
    if (kc <= 0) {
        kc = 16;
        for (const ZoneDesc& z : Z.host) kc = std::min(kc, visc_kc_default(z.lv[n]));
    }
    for (const ZoneDesc& z : Z.host) kc = std::min(kc, z.lv[n].nk - 2);
    unsigned nv = 0;
    const ZoneGrid V = Z.grid(
        [&](const ZoneDesc& z) { return visc_gx(z.lv[n]) * visc_gy(z.lv[n]) * (unsigned)((z.lv[n].nk - 2 + kc - 1) / kc); },
        nv);
    hipLaunchKernelGGL(k_visc_pk_z, dim3(nv), dim3(kTX * kTY), 0, s, V, n, recalc ? 0 : calc_pk, kc);
}

int visc_kc_default(const Level& L) {
    // planes per workgroup: enough workgroups (about 2048) to fill the chip
    const long tiles = (long)((L.ni - 2 + kTX - 2) / (kTX - 1)) * ((L.nj - 2 + kTY - 2) / (kTY - 1));
    return (int)std::max(1L, std::min(16L, tiles * (L.nk - 2) / 2048));
}

void launch_visc_pk(hipStream_t s, const Level& L, const Fields& F, const Consts& C, int n, int calc_pk, int kc) {
    if (kc <= 0) kc = visc_kc_default(L);
    kc = std::min(kc, L.nk - 2);
    const dim3 gv(visc_gx(L), visc_gy(L), (unsigned)((L.nk - 2 + kc - 1) / kc));
    hipLaunchKernelGGL(k_visc_pk, gv, dim3(kTX * kTY), 0, s, L, F, C, n, calc_pk, kc);
}

void launch_residual(hipStream_t s, const Level& L, const Fields& F, const Consts& C, int n, int recalc,
                     int calc_pk, int kc) {
    // two-wave workgroups on large levels (four per CU at two waves per
    // SIMD: 433 -> 424 us on the fine level of the 193x129x129 box); four
    // waves on small ones
    const int b = L.n >= 100000 ? 128 : 256;
    const dim3 grid((unsigned)((L.n + b - 1) / b));
    hipLaunchKernelGGL(k_residual, grid, dim3(b), 0, s, L, F, C, n, recalc);
    launch_visc_pk(s, L, F, C, n, recalc ? 0 : calc_pk, kc);
}

}  // namespace desalm
