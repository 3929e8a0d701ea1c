// k_sa.hip -- Spalart-Allmaras / DES eddy-viscosity update on the fine level.
//
// des_eddy (src/des_eddy.F90:132-297): vorticity measure once, then itk
// sub-iterations of explicit right-hand side -> three ADI line solves ->
// relaxed update -> boundary conditions, clamp and xnut = nu~ fv1.
// Kernels:
//   k_sa_prep   U/J (stored, rhs_contra_j) and wo (rhs_sa_source)
//   k_sa_rhs    rhs_sa_grdnu, rhs_sa_prod, rhs_sa_viscous, rhs_sa_flux,
//               rhs_sa_unst_visc, rhs_sa_implicit_source on the interior,
//               and the ADI operators' face coefficients and dtev J / 2
//               (LDS-staged tiles marching in k)
//   k_sa_sweep<D> the rhs_sa_adi_solver sweeps in i, j and k with every
//               row on chip (seg_sweep.h); D = 2 also applies the update of
//               des_eddy.F90:260-272
//   k_sa_bc<D>  rhs_sa_bcond faces, k_sa_clamp its nu~ floor and xnut
#include <algorithm>

#include "desalm_internal.h"
#include "seg_sweep.h"

namespace desalm {

namespace {

// Spalart-Allmaras constants (des_eddy.F90:63-73)
constexpr double cb1 = 0.1355, cb2 = 0.622, sig = 2.0 / 3.0, capa = 0.41, cw1 = 3.23906782, cw2 = 0.3,
                 cw3h = 64.0, cv1c = 357.911, cv2 = 5.0;

__device__ __forceinline__ void sa_prep_body(const Level& L, const Fields& F, unsigned bid, unsigned nblk) {
    const long t = (long)xcd_block(bid, nblk) * blockDim.x + threadIdx.x;
    if (t >= L.n) return;
    const int ni = L.ni, nj = L.nj, nk = L.nk;
    const int i = (int)(t % ni), j = (int)((t / ni) % nj), k = (int)(t / ((long)ni * nj));
    const long g = L.off + t;
    const bool bnd = i == 0 || i == ni - 1 || j == 0 || j == nj - 1 || k == 0 || k == nk - 1;
    // every load is issued before the first store: the neighbours of a
    // boundary node are replaced by the node itself (its wo is zero anyway)
    const long sI = bnd ? 0 : 1, sJ = bnd ? 0 : ni, sK = bnd ? 0 : (long)ni * nj;
    const double* U = F.q[1];
    const double* V = F.q[2];
    const double* W = F.q[3];
    const double u = U[g], v = V[g], w = W[g];
    const double c1 = F.csi[0][g], c2 = F.csi[1][g], c3 = F.csi[2][g];
    const double e1 = F.eta[0][g], e2 = F.eta[1][g], e3 = F.eta[2][g];
    const double z1 = F.zet[0][g], z2 = F.zet[1][g], z3 = F.zet[2][g];
    const double tmp = 1.0 / F.aj[g];
    const double uip = U[g + sI], uim = U[g - sI], vip = V[g + sI], vim = V[g - sI], wip = W[g + sI], wim = W[g - sI];
    const double ujp = U[g + sJ], ujm = U[g - sJ], vjp = V[g + sJ], vjm = V[g - sJ], wjp = W[g + sJ], wjm = W[g - sJ];
    const double ukp = U[g + sK], ukm = U[g - sK], vkp = V[g + sK], vkm = V[g - sK], wkp = W[g + sK], wkm = W[g - sK];
    F.ucn[0][t] = tmp * (c1 * u + c2 * v + c3 * w);
    F.ucn[1][t] = tmp * (e1 * u + e2 * v + e3 * w);
    F.ucn[2][t] = tmp * (z1 * u + z2 * v + z3 * w);
    if (bnd) {
        F.wo[t] = 0.0;
        return;
    }
    // rhs_sa_source.F90:44-93 (interior only)
    const double dc2 = 0.5 * L.dc, de2 = 0.5 * L.de, dz2 = 0.5 * L.dz;
    const double uc = dc2 * (uip - uim), vc = dc2 * (vip - vim), wc = dc2 * (wip - wim);
    const double ue = de2 * (ujp - ujm), ve = de2 * (vjp - vjm), we = de2 * (wjp - wjm);
    const double uz = dz2 * (ukp - ukm), vz = dz2 * (vkp - vkm), wz = dz2 * (wkp - wkm);
    const double u2 = uc * c2 + ue * e2 + uz * z2;
    const double u3 = uc * c3 + ue * e3 + uz * z3;
    const double v1 = vc * c1 + ve * e1 + vz * z1;
    const double v3 = vc * c3 + ve * e3 + vz * z3;
    const double w1 = wc * c1 + we * e1 + wz * z1;
    const double w2 = wc * c2 + we * e2 + wz * z2;
    const double w12 = 0.5 * (u2 - v1), w13 = 0.5 * (u3 - w1), w23 = 0.5 * (v3 - w2);
    F.wo[t] = 2.0 * (w12 * w12 + w13 * w13 + w23 * w23);
}

// ---- k_sa_rhs: rhs_sa_grdnu, rhs_sa_prod, rhs_sa_viscous, rhs_sa_flux,
// rhs_sa_unst_visc and rhs_sa_implicit_source on LDS-staged tiles ----
//
// Organised as k_visc_pk is: a workgroup of kQX x kQY threads owns a (kQX-1) x (kQY-1) tile of
// interior nodes plus a low halo column and row, and marches through kc
// planes.  nu~ of planes k-1, k, k+1, and the metrics, J and U/J, V/J of
// plane k, sit in LDS for the tile plus a high halo column and row.  Every
// +i and +j face (viscous flux, convective flux and the ADI coefficient
// av_D) is evaluated once and shared through LDS; the +k face is carried
// into the next plane.  The ADI coefficients are written on the faces the
// sweeps read (positions 0 .. n_D - 2 along D, interior across), dtj on the
// interior; rh5 on the interior (the sweeps never read its boundary).
// Operand order follows the reference routines, except for two
// rounding-level substitutions, both for speed (the kernel is bound by its
// FP64 arithmetic, not by memory):
// - the two powers of rhs_sa_prod.F90:47,84 are 1 / x^3 and cbrt(sqrt(y))
//   instead of pow(x, -3) and pow(y, 1/6), about a quarter of the kernel's
//   time;
// - every division is qdiv: a product with the Newton-refined hardware
//   reciprocal, about a tenth.  The ADI
// face coefficients are
//   av_D = |U/J face average| + (1/Re + nu~ face average)/sig * |grad csi_D|^2_face / (J + J')
// (rhs_sa_adi_solver.F90:53-75); nu~ does not change between the three
// sweeps of one sub-iteration.
constexpr int kQX = 32, kQY = 8;
constexpr int kQLX = kQX + 1, kQLY = kQY + 1, kQLN = kQLX * kQLY;
constexpr int kQHalo = kQLN - kQX * kQY;
constexpr int kQM = 10;  // csi 3, eta 3, zet 3, J

// a / b as a times the reciprocal of b: the hardware reciprocal refined by
// two Newton-Raphson steps r <- r (2 - b r) (quadratic convergence from the
// instruction's ~2^-40 estimate), then one product
__device__ __forceinline__ double qdiv(double a, double b) {
    double r = __builtin_amdgcn_rcp(b);
    r = fma(r, fma(-b, r, 1.0), r);
    r = fma(r, fma(-b, r, 1.0), r);
    return a * r;
}

struct QMet {
    double v[kQM];
};

__device__ __forceinline__ void qload_met(const Fields& F, long g, bool ok, QMet& m) {
#pragma unroll
    for (int c = 0; c < 3; c++) {
        m.v[c] = ok ? ld32(F.csi[c], (unsigned)(g)) : 0.0;
        m.v[3 + c] = ok ? ld32(F.eta[c], (unsigned)(g)) : 0.0;
        m.v[6 + c] = ok ? ld32(F.zet[c], (unsigned)(g)) : 0.0;
    }
    m.v[9] = ok ? ld32(F.aj, (unsigned)(g)) : 0.0;
}

// viscous flux of nu~ across the face node 0 -> node 0 + e_D
// (rhs_sa_viscous.F90:58-188); Q(a, b, c) is nu~ at node 0 + (a, b, c)
template <int D, typename Qf>
__device__ __forceinline__ double sa_visc_face_t(const Level& L, const Consts& C, Qf q, const QMet& m0,
                                                 const QMet& m1) {
    const double hc = L.dc, he = L.de, hz = L.dz;
    const double hc4 = 0.25 * hc, he4 = 0.25 * he, hz4 = 0.25 * hz;
    double vcs, vet, vze;
    if (D == 0) {
        vcs = hc * (q(1, 0, 0) - q(0, 0, 0));
        vet = he4 * (q(1, 1, 0) - q(1, -1, 0) + q(0, 1, 0) - q(0, -1, 0));
        vze = hz4 * (q(1, 0, 1) - q(1, 0, -1) + q(0, 0, 1) - q(0, 0, -1));
    } else if (D == 1) {
        vcs = hc4 * (q(1, 1, 0) - q(-1, 1, 0) + q(1, 0, 0) - q(-1, 0, 0));
        vet = he * (q(0, 1, 0) - q(0, 0, 0));
        vze = hz4 * (q(0, 1, 1) - q(0, 1, -1) + q(0, 0, 1) - q(0, 0, -1));
    } else {
        vcs = hc4 * (q(1, 0, 1) - q(-1, 0, 1) + q(1, 0, 0) - q(-1, 0, 0));
        vet = he4 * (q(0, 1, 1) - q(0, -1, 1) + q(0, 1, 0) - q(0, -1, 0));
        vze = hz * (q(0, 0, 1) - q(0, 0, 0));
    }
    const double a1 = m1.v[0] + m0.v[0], a2 = m1.v[1] + m0.v[1], a3 = m1.v[2] + m0.v[2];
    const double b1 = m1.v[3] + m0.v[3], b2 = m1.v[4] + m0.v[4], b3 = m1.v[5] + m0.v[5];
    const double c1 = m1.v[6] + m0.v[6], c2 = m1.v[7] + m0.v[7], c3 = m1.v[8] + m0.v[8];
    const double f1 = D == 0 ? a1 : D == 1 ? b1 : c1;
    const double f2 = D == 0 ? a2 : D == 1 ? b2 : c2;
    const double f3 = D == 0 ? a3 : D == 1 ? b3 : c3;
    const double cc1 = f1 * a1 + f2 * a2 + f3 * a3;
    const double cc2 = f1 * b1 + f2 * b2 + f3 * b3;
    const double cc3 = f1 * c1 + f2 * c2 + f3 * c3;
    const double q1 = D == 0 ? q(1, 0, 0) : D == 1 ? q(0, 1, 0) : q(0, 0, 1);
    const double q0 = q(0, 0, 0);
    const double ret1 = qdiv(C.rei + q1, sig), ret0 = qdiv(C.rei + q0, sig);
    const double vtj = qdiv(0.25 * (ret1 + ret0), m1.v[9] + m0.v[9]);
    return vtj * (cc1 * vcs + cc2 * vet + cc3 * vze);
}

// convective flux (rhs_sa_flux.F90:32-110) and ADI coefficient
// (rhs_sa_adi_solver.F90:53-75) of the face node 0 -> node 1 along D
template <int D>
__device__ __forceinline__ void sa_conv_av(const Consts& C, double q0, double q1, double u0, double u1,
                                           const QMet& m0, const QMet& m1, double& conv, double& av) {
    const double eigh = 0.5 * fabs(u1 + u0);
    conv = 0.5 * (q1 * u1 + q0 * u0 - eigh * (q1 - q0));
    const double gx = m1.v[3 * D] + m0.v[3 * D], gy = m1.v[3 * D + 1] + m0.v[3 * D + 1];
    const double gz = m1.v[3 * D + 2] + m0.v[3 * D + 2];
    const double ev = qdiv(C.rei + 0.5 * (q1 + q0), sig);
    const double ucna = 0.5 * fabs(u1 + u0);
    av = ucna + qdiv(ev * (gx * gx + gy * gy + gz * gz), m1.v[9] + m0.v[9]);
}

// tiles of the SA right-hand side along i and j
__host__ __device__ inline unsigned sa_gx(const Level& L) { return (unsigned)((L.ni - 2 + kQX - 2) / (kQX - 1)); }
__host__ __device__ inline unsigned sa_gy(const Level& L) { return (unsigned)((L.nj - 2 + kQY - 2) / (kQY - 1)); }

// bid: linear workgroup index over (i tile, j tile, k chunk), i fastest
__device__ __forceinline__ void sa_rhs_body(const Level& L, const Fields& F, const Consts& C, int kc, unsigned bid,
                                            unsigned nblk) {
    __shared__ double s_q[3][kQLN];      // nu~ of planes k-1, k, k+1 (slot k mod 3)
    __shared__ double s_m[kQM][kQLN];    // metrics, J of plane k
    __shared__ double s_u[2][kQLN];      // U/J, V/J of plane k
    __shared__ double s_fv[2][kQY][kQX];  // viscous flux of the +i, +j faces
    __shared__ double s_fc[2][kQY][kQX];  // convective flux of the +i, +j faces
    const int ni = L.ni, nj = L.nj, nk = L.nk;
    const unsigned gx = sa_gx(L), gy = sa_gy(L);
    const unsigned lb = xcd_block(bid, nblk);
    const int bx = (int)(lb % gx), by = (int)((lb / gx) % gy), bz = (int)(lb / (gx * gy));
    const int tid = (int)threadIdx.x;
    const int tx = tid % kQX, ty = tid / kQX;
    const int i0 = bx * (kQX - 1), j0 = by * (kQY - 1);
    const int kb = 1 + bz * kc, ke = min(kb + kc, nk - 1);
    const long sJ = ni, sK = (long)ni * nj;
    const int own = ty * kQLX + tx;
    const int hr = tid < kQLY ? tid : kQY, hcol = tid < kQLY ? kQX : tid - kQLY;
    const int halo = hr * kQLX + hcol;
    const bool has_halo = tid < kQHalo;
    const int i = i0 + tx, j = j0 + ty;
    const bool own_ok = i < ni && j < nj;
    const bool halo_ok = has_halo && i0 + hcol < ni && j0 + hr < nj;
    const long col = (long)i + sJ * j;                 // level-local
    const long hc = (long)(i0 + hcol) + sJ * (j0 + hr);
    const long o = L.off;
    const bool in_i = i <= ni - 2, in_j = j <= nj - 2;
    const bool out = tx >= 1 && ty >= 1 && in_i && in_j;
    const bool face_i = ty >= 1 && in_i && in_j;
    const bool face_j = tx >= 1 && in_i && in_j;
    const double* q5 = F.q[4];

    auto qmet_at = [&](int node) {
        QMet m;
#pragma unroll
        for (int c = 0; c < kQM; c++) m.v[c] = s_m[c][node];
        return m;
    };

    // prologue: nu~ of planes kb-1 .. kb+1; metrics, J, U/J, V/J of plane kb
    for (int k = kb - 1; k <= kb + 1; k++) {
        const int sl = k % 3;
        s_q[sl][own] = own_ok ? ld32(q5, (unsigned)(o + col + sK * k)) : 0.0;
        if (has_halo) s_q[sl][halo] = halo_ok ? ld32(q5, (unsigned)(o + hc + sK * k)) : 0.0;
    }
    QMet mcur, mh;
    qload_met(F, o + col + sK * kb, own_ok, mcur);
    qload_met(F, o + hc + sK * kb, halo_ok, mh);
#pragma unroll
    for (int c = 0; c < kQM; c++) {
        s_m[c][own] = mcur.v[c];
        if (has_halo) s_m[c][halo] = mh.v[c];
    }
#pragma unroll
    for (int d = 0; d < 2; d++) {
        s_u[d][own] = own_ok ? ld32(F.ucn[d], (unsigned)(col + sK * kb)) : 0.0;
        if (has_halo) s_u[d][halo] = halo_ok ? ld32(F.ucn[d], (unsigned)(hc + sK * kb)) : 0.0;
    }
    QMet mprev;
    qload_met(F, o + col + sK * (kb - 1), own_ok, mprev);
    double u2prev = own_ok ? ld32(F.ucn[2], (unsigned)(col + sK * (kb - 1))) : 0.0;
    double u2cur = own_ok ? ld32(F.ucn[2], (unsigned)(col + sK * kb)) : 0.0;
    __syncthreads();
    double fkv = 0.0, fkc = 0.0;  // fluxes of the face below the current plane
    if (out) {
        const int sl0 = (kb - 1) % 3, sl1 = kb % 3;
        auto q = [&](int a, int b, int c) { return s_q[c <= 0 ? sl0 : sl1][own + b * kQLX + a]; };
        fkv = sa_visc_face_t<2>(L, C, q, mprev, mcur);
        double av;
        sa_conv_av<2>(C, s_q[sl0][own], s_q[sl1][own], u2prev, u2cur, mprev, mcur, fkc, av);
        if (kb == 1) st32(F.av[2], (unsigned)(col), av);  // the face k = 1/2, position 0
    }
    for (int k = kb; k < ke; k++) {
        // next planes into registers
        QMet mnext, mhn;
        qload_met(F, o + col + sK * (k + 1), own_ok, mnext);
        qload_met(F, o + hc + sK * (k + 1), halo_ok, mhn);
        double un[2], uh[2];
#pragma unroll
        for (int d = 0; d < 2; d++) {
            un[d] = own_ok ? ld32(F.ucn[d], (unsigned)(col + sK * (k + 1))) : 0.0;
            uh[d] = halo_ok ? ld32(F.ucn[d], (unsigned)(hc + sK * (k + 1))) : 0.0;
        }
        const double u2next = own_ok ? ld32(F.ucn[2], (unsigned)(col + sK * (k + 1))) : 0.0;
        const bool more = k + 1 < ke;
        double qn_o = 0.0, qn_h = 0.0;
        if (more) {
            qn_o = own_ok ? ld32(q5, (unsigned)(o + col + sK * (k + 2))) : 0.0;
            qn_h = halo_ok ? ld32(q5, (unsigned)(o + hc + sK * (k + 2))) : 0.0;
        }
        const int slm = (k + 2) % 3, sl0 = k % 3, slp = (k + 1) % 3;
        auto q = [&](int a, int b, int c) {
            return s_q[c < 0 ? slm : c == 0 ? sl0 : slp][own + b * kQLX + a];
        };
        const long t = col + sK * k;  // level-local node

        if (face_i) {
            const QMet m1 = qmet_at(own + 1);
            s_fv[0][ty][tx] = sa_visc_face_t<0>(L, C, q, mcur, m1);
            double cv, av;
            sa_conv_av<0>(C, q(0, 0, 0), q(1, 0, 0), s_u[0][own], s_u[0][own + 1], mcur, m1, cv, av);
            s_fc[0][ty][tx] = cv;
            st32(F.av[0], (unsigned)(t), av);
        }
        if (face_j) {
            const QMet m1 = qmet_at(own + kQLX);
            s_fv[1][ty][tx] = sa_visc_face_t<1>(L, C, q, mcur, m1);
            double cv, av;
            sa_conv_av<1>(C, q(0, 0, 0), q(0, 1, 0), s_u[1][own], s_u[1][own + kQLX], mcur, m1, cv, av);
            s_fc[1][ty][tx] = cv;
            st32(F.av[1], (unsigned)(t), av);
        }
        double fkv1 = 0.0, fkc1 = 0.0;
        double dnu = 0.0;
        if (out) {
            fkv1 = sa_visc_face_t<2>(L, C, q, mcur, mnext);
            double av;
            sa_conv_av<2>(C, q(0, 0, 0), q(0, 0, 1), u2cur, u2next, mcur, mnext, fkc1, av);
            st32(F.av[2], (unsigned)(t), av);
            // gradient of nu~ (rhs_sa_grdnu.F90:22-60)
            const double dc2 = 0.5 * L.dc, de2 = 0.5 * L.de, dz2 = 0.5 * L.dz;
            const double qip = qdiv(q(1, 0, 0), s_m[9][own + 1]), qim = qdiv(q(-1, 0, 0), s_m[9][own - 1]);
            const double qjp = qdiv(q(0, 1, 0), s_m[9][own + kQLX]), qjm = qdiv(q(0, -1, 0), s_m[9][own - kQLX]);
            const double qkp = qdiv(q(0, 0, 1), mnext.v[9]), qkm = qdiv(q(0, 0, -1), mprev.v[9]);
            double d[3];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                d[c] = dc2 * (qip * s_m[c][own + 1] - qim * s_m[c][own - 1]);
                d[c] = d[c] + de2 * (qjp * s_m[3 + c][own + kQLX] - qjm * s_m[3 + c][own - kQLX]);
                d[c] = d[c] + dz2 * (qkp * mnext.v[6 + c] - qkm * mprev.v[6 + c]);
            }
            const double aj2 = mcur.v[9] * mcur.v[9];
            dnu = aj2 * (d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        }
        __syncthreads();
        if (out) {
            // every point input loaded before the first use, so that the
            // loads are in flight together
            const unsigned g = (unsigned)(o + t);
            const double wdv = ld32(F.wd, g), wov = ld32(F.wo, (unsigned)t);
            const double qn5 = ld32(F.qn[4], g), qnm15 = ld32(F.qnm1[4], g), dtev = ld32(F.dtev, g);
            const double aj = mcur.v[9];
            // production, destruction, cb2 diffusion (rhs_sa_prod.F90:44-100)
            double psa;
            if (wdv == 0.0) {
                psa = 0.0;
            } else {
                const double qv = q(0, 0, 0);
                const double chi = qdiv(qv, C.rei);
                const double chic = chi * chi * chi;
                const double fv1 = qdiv(chic, chic + cv1c);
                const double xf = 1.0 + qdiv(chi, cv2);
                const double fv2 = qdiv(1.0, xf * xf * xf);
                const double fv3 = qdiv((1.0 + chi * fv1) * (1.0 - fv2), chi);
                const double vtcm = sqrt(2.0 * wov);
                const double cawd = capa * capa * wdv * wdv;
                const double vtc_org = fv3 * vtcm + qdiv(qv * fv2, cawd);
                const double vtc = fmax(vtc_org, 0.00001);
                psa = cb1 * vtc * qv;
                const double rr = qdiv(qv, vtc * cawd);
                const double rr6 = rr * rr * rr * rr * rr * rr;
                const double grr = fmax(0.0, rr + cw2 * (rr6 - rr));
                const double grr6 = grr * grr * grr * grr * grr * grr;
                const double fw = grr * cbrt(sqrt(qdiv(1.0 + cw3h, grr6 + cw3h)));
                const double vwd = qdiv(qv, wdv);
                psa = psa - cw1 * fw * vwd * vwd;
                psa = psa + qdiv(cb2 * dnu, sig);
            }
            double visc = s_fv[0][ty][tx] - s_fv[0][ty][tx - 1];
            visc = visc + s_fv[1][ty][tx] - s_fv[1][ty - 1][tx];
            visc = visc + fkv1 - fkv;
            double r = s_fc[0][ty][tx] - s_fc[0][ty][tx - 1];
            r = r + s_fc[1][ty][tx] - s_fc[1][ty - 1][tx];
            r = r + fkc1 - fkc;
            const double dtjb = 2.0 * C.delti * aj;
            const double dq5dt = C.e_source * qdiv(3.0 * q(0, 0, 0) - 4.0 * qn5 + qnm15, dtjb);
            r = r + dq5dt - visc - qdiv(psa, aj);
            st32(F.rh5, (unsigned)(t), -dtev * aj * r);
            st32(F.dtj, (unsigned)(t), 0.5 * dtev * aj);
        }
        // plane k+1 into LDS: no barrier needed before it, as the part above
        // reads only the face fluxes and nu~ of plane k (slot k mod 3), and
        // the next plane's nu~ goes to the slot of plane k-1
#pragma unroll
        for (int c = 0; c < kQM; c++) {
            s_m[c][own] = mnext.v[c];
            if (has_halo) s_m[c][halo] = mhn.v[c];
        }
#pragma unroll
        for (int d = 0; d < 2; d++) {
            s_u[d][own] = un[d];
            if (has_halo) s_u[d][halo] = uh[d];
        }
        if (more) {
            s_q[slm][own] = qn_o;
            if (has_halo) s_q[slm][halo] = qn_h;
        }
        mprev = mcur;
        mcur = mnext;
        u2cur = u2next;
        fkv = fkv1;
        fkc = fkc1;
        __syncthreads();
    }
}

// rhs_sa_bcond faces (rhs_sa_bcond.F90:40-269); faces of direction D, both sides
template <int D>
__device__ __forceinline__ void sa_bc_body(const Level& L, const Fields& F, const Consts& C, unsigned bid) {
    const int ni = L.ni, nj = L.nj, nk = L.nk;
    const long sI = 1, sJ = ni, sK = (long)ni * nj;
    const long t = (long)bid * blockDim.x + threadIdx.x;
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
        case 6: case 7: q5[lo] = q5[base + (long)(nd - 2) * sD]; break;  // wraps in both
        case 5: {  // characteristic outlet: extrapolated on the grid ratio (k faces: ratio 0)
            const double r = D < 2 ? F.rat[b0][t] : 0.0;
            q5[lo] = (1.0 + r) * q5[lo + sD] - r * q5[lo + 2 * sD];
            break;
        }
        default: break;
    }
    switch (C.btype[b1]) {
        case 1: q5[hi] = 0.0; break;
        case 2: case 3: q5[hi] = C.sa[b1][0] * q5[hi - sD] + C.sb[b1][0] * q5[hi - 2 * sD]; break;
        case 6: case 7: q5[hi] = q5[base + sD]; break;
        case 8:
            if (D == 0) q5[hi] = q5[hi - sD];
            break;
        case 5: {
            const double r = D < 2 ? F.rat[b1][t] : 0.0;
            q5[hi] = (1.0 + r) * q5[hi - sD] - r * q5[hi - 2 * sD];
            break;
        }
        default: break;
    }
}

// nu~ floor and xnut = nu~ fv1 on every node (rhs_sa_bcond.F90:279-291)
__device__ __forceinline__ void sa_clamp_body(const Level& L, const Fields& F, const Consts& C, unsigned bid) {
    const long t = (long)bid * blockDim.x + threadIdx.x;
    if (t >= L.n) return;
    const long g = L.off + t;
    const double q = fmax(1.0e-07, F.q[4][g]);
    F.q[4][g] = q;
    const double chi = q / C.rei;
    const double chic = chi * chi * chi;
    const double fv1 = chic / (chic + cv1c);
    F.xnut[g] = q * fv1;
}

// One ADI sweep of rhs_sa_adi_solver (src/rhs_sa_adi_solver.F90:44-244) on
// every interior line of direction D with the rows on chip (seg_sweep.h):
//   a = -dtj (U_(s-1) + av_(s-1/2)),  b = 1 + dtj (av_(s-1/2) + av_(s+1/2)),
//   c = dtj (U_(s+1) - av_(s+1/2)),   right-hand side rh5, solved in place
// (a line's rows are read before and written after the workgroup's barriers).
// D = 2 also applies the relaxed update of des_eddy.F90:260-272:
//   q5 += (-q5 + qold + x) / (1 + 1.5 e_source dtev / delti), qold = q5 here.
template <int D, int R>
__device__ __forceinline__ void sa_sweep_body(const Level& L, const Fields& F, const Consts& C, const SegShape& sh,
                                              unsigned bid) {
    extern __shared__ double smem_sa[];
    const int P = sh.P, TI = sh.TI, n = sh.n;
    const int tid = (int)threadIdx.x;
    const int li = D == 0 ? tid / P : tid % TI;
    const int ps = D == 0 ? tid % P : tid / TI;
    const long line0 = (long)bid * TI;
    const long t = line0 + li;
    const bool act = t < sh.nlines;
    const int ni = L.ni, nj = L.nj;
    long base = 0, sD;
    if (D == 0) {
        sD = 1;
        if (act) base = (long)ni * ((1 + t % (nj - 2)) + (long)nj * (1 + t / (nj - 2)));
    } else if (D == 1) {
        sD = ni;
        if (act) base = (1 + t % (ni - 2)) + (long)ni * nj * (1 + t / (ni - 2));
    } else {
        sD = (long)ni * nj;
        if (act) base = (1 + t % (ni - 2)) + (long)ni * (1 + t / (ni - 2));
    }
    const int s0 = ps * R;
    const int nr = act ? max(0, min(R, n - s0)) : 0;
    const double* __restrict__ U = F.ucn[D];
    const double* __restrict__ av = F.av[D];
    const double* __restrict__ dtj = F.dtj;
    double* __restrict__ rh = F.rh5;
    double a[R], b[R], c[R], d[R][1];
    double a_next;
    if (D == 0) {
        // i lines: the segment's rows and their neighbours are one run of
        // consecutive nodes per array, loaded as such (16-byte loads; lanes
        // are R rows apart).  The run of a partial last segment (or an idle
        // thread) reaches past the line's end into the next line or plane of
        // the level; those values are never used.  Run entry e is node
        // position s0 + e (av, U) or s0 + 1 + e (dtj, rh).
        double avw[R + 1], uw[R + 2], dtw[R + 1], rhw[R];
#pragma unroll
        for (int e = 0; e < R + 2; e++) {
            const long tl = base + s0 + e;
            uw[e] = U[tl];
            if (e <= R) avw[e] = av[tl];
            if (e <= R) dtw[e] = dtj[tl + 1];
            if (e < R) rhw[e] = rh[tl + 1];
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
            const double avm = avw[r], avp = avw[r + 1], dt = dtw[r];
            const double ar = -dt * (uw[r] + avm);
            const double br = 1.0 + dt * (avm + avp);
            const double cr = dt * (uw[r + 2] - avp);
            const bool in = r < nr;
            a[r] = in ? ar : 0.0;
            b[r] = in ? br : 1.0;
            c[r] = in ? cr : 0.0;
            d[r][0] = in ? rhw[r] : 0.0;
        }
        const double an = -dtw[R] * (uw[R] + avw[R]);
        a_next = (nr == R && s0 + R < n) ? an : 0.0;
    } else {
        // every row is loaded without a branch, so that the loads of all rows
        // can be in flight together: rows past the line's end (and every row of
        // an idle thread) read the line's last row and are discarded
#pragma unroll
        for (int r = 0; r < R; r++) {
            const long tl = base + (long)(min(s0 + r, n - 1) + 1) * sD;
            const double avm = av[tl - sD], avp = av[tl], dt = dtj[tl];
            const double ar = -dt * (U[tl - sD] + avm);
            const double br = 1.0 + dt * (avm + avp);
            const double cr = dt * (U[tl + sD] - avp);
            const double dr = rh[tl];
            const bool in = r < nr;
            a[r] = in ? ar : 0.0;
            b[r] = in ? br : 1.0;
            c[r] = in ? cr : 0.0;
            d[r][0] = in ? dr : 0.0;
        }
        {
            const long tl = base + (long)(min(s0 + nr, n - 1) + 1) * sD;
            const double an = -dtj[tl] * (U[tl - sD] + av[tl - sD]);
            a_next = (nr > 0 && s0 + nr < n) ? an : 0.0;
        }
    }
    unsigned piv = 0;
    int zrow = 0;
    seg_solve<R, 1, double>(a, b, c, d, a_next, nr, n, s0, smem_sa, P, TI, sh.nlines, line0, li, ps, piv, zrow);
    if (D == 0 && nr == R) {
        // a whole segment of an i line: one run of consecutive nodes
#pragma unroll
        for (int r = 0; r < R; r++) rh[base + s0 + r + 1] = d[r][0];
    } else if (D < 2) {
#pragma unroll
        for (int r = 0; r < R; r++)
            if (r < nr) rh[base + (long)(s0 + r + 1) * sD] = d[r][0];
    } else {
        // nu~ and dtev of every row are loaded, and every new nu~ formed,
        // before the first store: the loads are in flight together and the
        // stores wait for nothing
        double qv[R], dv[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const long g = L.off + base + (long)(min(s0 + r, n - 1) + 1) * sD;
            qv[r] = F.q[4][g];
            dv[r] = F.dtev[g];
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
            const double qe5 = qv[r] + d[r][0];
            const double dum = 1.0 + 1.5 * C.e_source * dv[r] / C.delti;
            qv[r] = qv[r] + (-qv[r] + qe5) / dum;
        }
#pragma unroll
        for (int r = 0; r < R; r++)
            if (r < nr) F.q[4][L.off + base + (long)(s0 + r + 1) * sD] = qv[r];
    }
    if (piv) atomicAdd(F.pivots, (unsigned long long)piv);
    if (zrow) atomicMin(F.zero_pivot, (unsigned)zrow);
}

__global__ void __launch_bounds__(256) k_sa_prep(Level L, Fields F) { sa_prep_body(L, F, blockIdx.x, gridDim.x); }
__global__ void __launch_bounds__(256) k_sa_prep_z(ZoneGrid G) {
    const ZoneBlock b = zone_block(G);
    sa_prep_body(G.z[b.z].lv[1], G.z[b.z].F, b.bid, b.nblk);
}
// three waves per SIMD: fits in 168 VGPRs without spilling
__global__ void __launch_bounds__(kQX * kQY, 3) k_sa_rhs(Level L, Fields F, Consts C, int kc) {
    sa_rhs_body(L, F, C, kc, blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z),
                gridDim.x * gridDim.y * gridDim.z);
}
__global__ void __launch_bounds__(kQX * kQY, 3) k_sa_rhs_z(ZoneGrid G, int kc) {
    const ZoneBlock b = zone_block(G);
    const ZoneDesc& Z = G.z[b.z];
    sa_rhs_body(Z.lv[1], Z.F, Z.C, kc, b.bid, b.nblk);
}
template <int D>
__global__ void __launch_bounds__(256) k_sa_bc(Level L, Fields F, Consts C) {
    sa_bc_body<D>(L, F, C, blockIdx.x);
}
template <int D>
__global__ void __launch_bounds__(256) k_sa_bc_z(ZoneGrid G) {
    const ZoneBlock b = zone_block(G);
    const ZoneDesc& Z = G.z[b.z];
    sa_bc_body<D>(Z.lv[1], Z.F, Z.C, b.bid);
}
__global__ void __launch_bounds__(256) k_sa_clamp(Level L, Fields F, Consts C) { sa_clamp_body(L, F, C, blockIdx.x); }
__global__ void __launch_bounds__(256) k_sa_clamp_z(ZoneGrid G) {
    const ZoneBlock b = zone_block(G);
    const ZoneDesc& Z = G.z[b.z];
    sa_clamp_body(Z.lv[1], Z.F, Z.C, b.bid);
}
template <int D, int R>
__global__ void __launch_bounds__(512) k_sa_sweep(Level L, Fields F, Consts C, SegShape sh) {
    sa_sweep_body<D, R>(L, F, C, sh, xcd_block(blockIdx.x, gridDim.x));
}
template <int D, int R>
__global__ void __launch_bounds__(512) k_sa_sweep_z(ZoneGrid G, int P, int TI) {
    const ZoneBlock b = zone_block(G);
    const ZoneDesc& Z = G.z[b.z];
    sa_sweep_body<D, R>(Z.lv[1], Z.F, Z.C, seg_lines(Z.lv[1], D, P, TI), xcd_block(b.bid, b.nblk));
}

}  // namespace

int sa_kc_default(const Level& L) {
    const long tiles = (long)((L.ni - 2 + kQX - 2) / (kQX - 1)) * ((L.nj - 2 + kQY - 2) / (kQY - 1));
    return (int)std::max(1L, std::min(16L, tiles * (L.nk - 2) / 2048));
}

void launch_sa_prep(hipStream_t s, const Level& L, const Fields& F) {
    const int b = 256;
    hipLaunchKernelGGL(k_sa_prep, dim3((unsigned)((L.n + b - 1) / b)), dim3(b), 0, s, L, F);
}

void launch_sa_rhs(hipStream_t s, const Level& L, const Fields& F, const Consts& C, int kc) {
    if (kc <= 0) kc = sa_kc_default(L);
    kc = std::min(kc, L.nk - 2);
    const dim3 gq(sa_gx(L), sa_gy(L), (unsigned)((L.nk - 2 + kc - 1) / kc));
    hipLaunchKernelGGL(k_sa_rhs, gq, dim3(kQX * kQY), 0, s, L, F, C, kc);
}

// the ADI sweep of direction D over the interior lines of the fine level
template <int D>
static void sa_sweep(hipStream_t s, const Level& L, const Fields& F, const Consts& C) {
    const int n = D == 0 ? L.ni - 2 : D == 1 ? L.nj - 2 : L.nk - 2;
    const long nl = D == 0 ? (long)(L.nj - 2) * (L.nk - 2) : D == 1 ? (long)(L.ni - 2) * (L.nk - 2)
                                                                    : (long)(L.ni - 2) * (L.nj - 2);
    // eight rows per segment, four on lines of at most 64 rows (the i lines
    // load their rows as runs, so their segment length sets no access
    // stride; the DAF and rs i sweeps, with more values per row, run faster
    // at four)
    const int R = n <= 64 ? 4 : 8;
    const SegShape sh = seg_shape(n, nl, R, D == 0);
    const unsigned nb = (unsigned)((sh.nlines + sh.TI - 1) / sh.TI);
    const size_t lds = seg_lds_bytes(sh, 1);
    if (R == 4)
        hipLaunchKernelGGL((k_sa_sweep<D, 4>), dim3(nb), dim3(sh.TI * sh.P), lds, s, L, F, C, sh);
    else
        hipLaunchKernelGGL((k_sa_sweep<D, 8>), dim3(nb), dim3(sh.TI * sh.P), lds, s, L, F, C, sh);
}

void launch_des_eddy(hipStream_t s, const Level& L, const Fields& F, const Consts& C, int itk, int kc) {
    const int b = 256;
    const unsigned gn = (unsigned)((L.n + b - 1) / b);
    launch_sa_prep(s, L, F);
    const unsigned fi = (unsigned)(((long)L.nj * L.nk + b - 1) / b);
    const unsigned fj = (unsigned)(((long)L.ni * L.nk + b - 1) / b);
    const unsigned fk = (unsigned)(((long)L.ni * L.nj + b - 1) / b);
    for (int it = 0; it < itk; it++) {
        launch_sa_rhs(s, L, F, C, kc);
        sa_sweep<0>(s, L, F, C);
        sa_sweep<1>(s, L, F, C);
        sa_sweep<2>(s, L, F, C);
        hipLaunchKernelGGL(k_sa_bc<0>, dim3(fi), dim3(b), 0, s, L, F, C);
        hipLaunchKernelGGL(k_sa_bc<1>, dim3(fj), dim3(b), 0, s, L, F, C);
        hipLaunchKernelGGL(k_sa_bc<2>, dim3(fk), dim3(b), 0, s, L, F, C);
        hipLaunchKernelGGL(k_sa_clamp, dim3(gn), dim3(b), 0, s, L, F, C);
    }
}

// the ADI sweep of direction D over every zone of a group in one launch,
// with the segment count of the longest line
template <int D>
static void sa_sweep_z(hipStream_t s, const ZoneList& Z) {
    int len = 0;
    for (const ZoneDesc& z : Z.host) len = std::max(len, seg_lines(z.lv[1], D, 1, 1).n);
    const int R = len <= 64 ? 4 : 8;
    const SegShape sh = seg_shape(len, 1, R, D == 0);
    unsigned nb = 0;
    const ZoneGrid G = Z.grid(
        [&](const ZoneDesc& z) { return (unsigned)((seg_lines(z.lv[1], D, sh.P, sh.TI).nlines + sh.TI - 1) / sh.TI); },
        nb);
    const size_t lds = seg_lds_bytes(sh, 1);
    if (R == 4)
        hipLaunchKernelGGL((k_sa_sweep_z<D, 4>), dim3(nb), dim3(sh.TI * sh.P), lds, s, G, sh.P, sh.TI);
    else
        hipLaunchKernelGGL((k_sa_sweep_z<D, 8>), dim3(nb), dim3(sh.TI * sh.P), lds, s, G, sh.P, sh.TI);
}

// des_eddy on the fine level of every zone of a group, one launch per kernel
void launch_des_eddy_z(hipStream_t s, const ZoneList& Z, int itk, int kc) {
    const int b = 256;
    auto nodes = [&](const ZoneDesc& z) { return (unsigned)((z.lv[1].n + b - 1) / b); };
    unsigned gn = 0, fi = 0, fj = 0, fk = 0, gq = 0;
    const ZoneGrid Gn = Z.grid(nodes, gn);
    const ZoneGrid Gi = Z.grid([&](const ZoneDesc& z) { return (unsigned)(((long)z.lv[1].nj * z.lv[1].nk + b - 1) / b); }, fi);
    const ZoneGrid Gj = Z.grid([&](const ZoneDesc& z) { return (unsigned)(((long)z.lv[1].ni * z.lv[1].nk + b - 1) / b); }, fj);
    const ZoneGrid Gk = Z.grid([&](const ZoneDesc& z) { return (unsigned)(((long)z.lv[1].ni * z.lv[1].nj + b - 1) / b); }, fk);
    if (kc <= 0) {
        kc = 16;
        for (const ZoneDesc& z : Z.host) kc = std::min(kc, sa_kc_default(z.lv[1]));
    }
    for (const ZoneDesc& z : Z.host) kc = std::min(kc, z.lv[1].nk - 2);
    const ZoneGrid Gq = Z.grid(
        [&](const ZoneDesc& z) { return sa_gx(z.lv[1]) * sa_gy(z.lv[1]) * (unsigned)((z.lv[1].nk - 2 + kc - 1) / kc); },
        gq);
    hipLaunchKernelGGL(k_sa_prep_z, dim3(gn), dim3(b), 0, s, Gn);
    for (int it = 0; it < itk; it++) {
        hipLaunchKernelGGL(k_sa_rhs_z, dim3(gq), dim3(kQX * kQY), 0, s, Gq, kc);
        sa_sweep_z<0>(s, Z);
        sa_sweep_z<1>(s, Z);
        sa_sweep_z<2>(s, Z);
        hipLaunchKernelGGL(k_sa_bc_z<0>, dim3(fi), dim3(b), 0, s, Gi);
        hipLaunchKernelGGL(k_sa_bc_z<1>, dim3(fj), dim3(b), 0, s, Gj);
        hipLaunchKernelGGL(k_sa_bc_z<2>, dim3(fk), dim3(b), 0, s, Gk);
        hipLaunchKernelGGL(k_sa_clamp_z, dim3(gn), dim3(b), 0, s, Gn);
    }
}

}  // namespace desalm
