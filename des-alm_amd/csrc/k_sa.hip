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
//   k_sa_adi_i  rhs_sa_adi_solver sweep in i (rows staged through LDS)
//   k_sa_adi<D> the sweeps in j and k (D = 2 also applies the update of
//               des_eddy.F90:260-272)
//   k_sa_bc<D>  rhs_sa_bcond faces, k_sa_clamp its nu~ floor and xnut
#include <algorithm>

#include "desalm_internal.h"
#include "line_tile.h"

namespace desalm {

namespace {

// Spalart-Allmaras constants (des_eddy.F90:63-73)
constexpr double cb1 = 0.1355, cb2 = 0.622, sig = 2.0 / 3.0, capa = 0.41, cw1 = 3.23906782, cw2 = 0.3,
                 cw3h = 64.0, cv1c = 357.911, cv2 = 5.0;

__global__ void k_sa_prep(Level L, Fields F) {
    const long t = (long)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
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
        m.v[c] = ok ? F.csi[c][g] : 0.0;
        m.v[3 + c] = ok ? F.eta[c][g] : 0.0;
        m.v[6 + c] = ok ? F.zet[c][g] : 0.0;
    }
    m.v[9] = ok ? F.aj[g] : 0.0;
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

__global__ void __launch_bounds__(kQX * kQY) k_sa_rhs(Level L, Fields F, Consts C, int kc) {
    __shared__ double s_q[3][kQLN];      // nu~ of planes k-1, k, k+1 (slot k mod 3)
    __shared__ double s_m[kQM][kQLN];    // metrics, J of plane k
    __shared__ double s_u[2][kQLN];      // U/J, V/J of plane k
    __shared__ double s_fv[2][kQY][kQX];  // viscous flux of the +i, +j faces
    __shared__ double s_fc[2][kQY][kQX];  // convective flux of the +i, +j faces
    const int ni = L.ni, nj = L.nj, nk = L.nk;
    const unsigned gx = gridDim.x, gy = gridDim.y;
    const unsigned lb = xcd_block(blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z), gx * gy * gridDim.z);
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
        s_q[sl][own] = own_ok ? q5[o + col + sK * k] : 0.0;
        if (has_halo) s_q[sl][halo] = halo_ok ? q5[o + hc + sK * k] : 0.0;
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
        s_u[d][own] = own_ok ? F.ucn[d][col + sK * kb] : 0.0;
        if (has_halo) s_u[d][halo] = halo_ok ? F.ucn[d][hc + sK * kb] : 0.0;
    }
    QMet mprev;
    qload_met(F, o + col + sK * (kb - 1), own_ok, mprev);
    double u2prev = own_ok ? F.ucn[2][col + sK * (kb - 1)] : 0.0;
    double u2cur = own_ok ? F.ucn[2][col + sK * kb] : 0.0;
    __syncthreads();
    double fkv = 0.0, fkc = 0.0;  // fluxes of the face below the current plane
    if (out) {
        const int sl0 = (kb - 1) % 3, sl1 = kb % 3;
        auto q = [&](int a, int b, int c) { return s_q[c <= 0 ? sl0 : sl1][own + b * kQLX + a]; };
        fkv = sa_visc_face_t<2>(L, C, q, mprev, mcur);
        double av;
        sa_conv_av<2>(C, s_q[sl0][own], s_q[sl1][own], u2prev, u2cur, mprev, mcur, fkc, av);
        if (kb == 1) F.av[2][col] = av;  // the face k = 1/2, position 0
    }
    for (int k = kb; k < ke; k++) {
        // next planes into registers
        QMet mnext, mhn;
        qload_met(F, o + col + sK * (k + 1), own_ok, mnext);
        qload_met(F, o + hc + sK * (k + 1), halo_ok, mhn);
        double un[2], uh[2];
#pragma unroll
        for (int d = 0; d < 2; d++) {
            un[d] = own_ok ? F.ucn[d][col + sK * (k + 1)] : 0.0;
            uh[d] = halo_ok ? F.ucn[d][hc + sK * (k + 1)] : 0.0;
        }
        const double u2next = own_ok ? F.ucn[2][col + sK * (k + 1)] : 0.0;
        const bool more = k + 1 < ke;
        double qn_o = 0.0, qn_h = 0.0;
        if (more) {
            qn_o = own_ok ? q5[o + col + sK * (k + 2)] : 0.0;
            qn_h = halo_ok ? q5[o + hc + sK * (k + 2)] : 0.0;
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
            F.av[0][t] = av;
        }
        if (face_j) {
            const QMet m1 = qmet_at(own + kQLX);
            s_fv[1][ty][tx] = sa_visc_face_t<1>(L, C, q, mcur, m1);
            double cv, av;
            sa_conv_av<1>(C, q(0, 0, 0), q(0, 1, 0), s_u[1][own], s_u[1][own + kQLX], mcur, m1, cv, av);
            s_fc[1][ty][tx] = cv;
            F.av[1][t] = av;
        }
        double fkv1 = 0.0, fkc1 = 0.0;
        double dnu = 0.0;
        if (out) {
            fkv1 = sa_visc_face_t<2>(L, C, q, mcur, mnext);
            double av;
            sa_conv_av<2>(C, q(0, 0, 0), q(0, 0, 1), u2cur, u2next, mcur, mnext, fkc1, av);
            F.av[2][t] = av;
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
            const long g = o + t;
            const double aj = mcur.v[9];
            // production, destruction, cb2 diffusion (rhs_sa_prod.F90:44-100)
            double psa;
            const double wdv = F.wd[g];
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
                const double vtcm = sqrt(2.0 * F.wo[t]);
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
            const double dq5dt = C.e_source * qdiv(3.0 * q(0, 0, 0) - 4.0 * F.qn[4][g] + F.qnm1[4][g], dtjb);
            r = r + dq5dt - visc - qdiv(psa, aj);
            const double dtev = F.dtev[g];
            F.rh5[t] = -dtev * aj * r;
            F.dtj[t] = 0.5 * dtev * aj;
        }
        __syncthreads();
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

// Two-way form of the j and k sweeps: each wave solves 32 lines, lanes
// 0-31 eliminating the upper halves of the lines downwards (rows 0 .. m-1,
// the Thomas recurrence of k_sa_adi), lanes 32-63 the lower halves upwards
// (rows len-1 .. m+1, the mirrored recurrence e' = aw / den,
// f' = (aq - ae f') / den).  The two meet at row m through a lane shuffle,
// then substitute outwards.  Twice the waves and half the serial chain of
// the one-way sweep, which runs one wave per CU at 129^3; rows below m are
// eliminated in the other order, a rounding-level difference.  Rows where
// LAPACK SGTSV would pivot are counted on the upper halves only.
constexpr int kPF2 = 16;  // rows per prefetch group: 16 measured best among 4, 8, 16

template <int D>
__global__ void __launch_bounds__(128) k_sa_adi2(Level L, Fields F, Consts C) {
    const int ni = L.ni, nj = L.nj, nk = L.nk;
    const int lane = (int)(threadIdx.x & 63), wave = (int)(threadIdx.x >> 6);
    const bool top = lane < 32;
    const long nlines = D == 1 ? (long)(ni - 2) * (nk - 2) : (long)(ni - 2) * (nj - 2);
    const long t = (long)blockIdx.x * 64 + wave * 32 + (lane & 31);
    const bool act = t < nlines;
    const long sD = D == 1 ? (long)ni : (long)ni * nj;
    const int len = D == 1 ? nj - 2 : nk - 2;
    long base = 0;
    if (act) base = D == 1 ? (1 + t % (ni - 2)) + (long)ni * nj * (1 + t / (ni - 2))
                           : (1 + t % (ni - 2)) + (long)ni * (1 + t / (ni - 2));
    const double* __restrict__ U = F.ucn[D];
    const double* __restrict__ av = F.av[D];
    const double* __restrict__ dtj = F.dtj;
    double* __restrict__ rh = F.rh5;
    double* __restrict__ cp = F.cp[0];
    const int m = len / 2;
    auto node = [&](int pos) { return base + (long)pos * sD; };  // line position 0 .. len + 1
    // upper half: carries av(s), U(s), U(s+1); lower half: av(s+1), U(s+1), U(s+2)
    double a_c = 0.0, u_m = 0.0, u_0 = 0.0;
    if (act) {
        if (top) { a_c = av[node(0)]; u_m = U[node(0)]; u_0 = U[node(1)]; }
        else { a_c = av[node(len)]; u_m = U[node(len + 1)]; u_0 = U[node(len)]; }
    }
    double cprev = 0.0, dprev = 0.0, denprev = 1.0;
    unsigned piv = 0;
    const int nrows = top ? m : len - 1 - m;
    for (int g0 = 0; g0 < nrows; g0 += kPF2) {
        double la[kPF2], lu[kPF2], ld[kPF2], lq[kPF2];
#pragma unroll
        for (int u = 0; u < kPF2; u++) {
            const int r = g0 + u;
            const int srow = top ? r : len - 1 - r;
            const bool ok = act && r < nrows;
            const long p = node(srow + 1);
            la[u] = ok ? av[top ? node(srow + 1) : node(srow)] : 0.0;
            lu[u] = ok ? U[top ? node(srow + 2) : node(srow)] : 0.0;
            ld[u] = ok ? dtj[p] : 0.0;
            lq[u] = ok ? rh[p] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kPF2; u++) {
            const int r = g0 + u;
            if (r < nrows) {
                const int srow = top ? r : len - 1 - r;
                double avm, avp, um, up;
                if (top) { avm = a_c; avp = la[u]; um = u_m; up = lu[u]; }
                else { avp = a_c; avm = la[u]; up = u_m; um = lu[u]; }
                const double dt = ld[u];
                const double aw = -dt * (um + avm);
                const double ap = 1.0 + dt * (avm + avp);
                const double ae = dt * (up - avp);
                const double aq = lq[u];
                double den;
                if (r == 0) {
                    den = ap;
                } else {
                    if (top && fabs(denprev) < fabs(aw)) piv++;
                    den = top ? ap - aw * cprev : ap - ae * cprev;
                }
                const double inv = 1.0 / den;
                if (top) {
                    cprev = ae * inv;
                    dprev = (r == 0) ? aq * inv : (aq - aw * dprev) * inv;
                } else {
                    cprev = aw * inv;
                    dprev = (r == 0) ? aq * inv : (aq - ae * dprev) * inv;
                }
                denprev = den;
                if (act) {
                    const long p = node(srow + 1);
                    cp[p] = cprev;
                    rh[p] = dprev;
                }
                // carry the shared face and velocities to the next row
                if (top) { a_c = avp; u_m = u_0; u_0 = up; }
                else { a_c = avm; u_m = u_0; u_0 = um; }
            }
        }
    }
    // the middle row m, solved by the upper lane with the lower half's e', f'
    const double e_lo = __shfl(cprev, (lane & 31) + 32, 64);
    const double f_lo = __shfl(dprev, (lane & 31) + 32, 64);
    double xm = 0.0;
    if (top && act) {
        const long p = node(m + 1);
        const double avm = a_c, avp = av[node(m + 1)], um = u_m, up = U[node(m + 2)];
        const double dt = dtj[p], aq = rh[p];
        const double aw = -dt * (um + avm);
        const double ap = 1.0 + dt * (avm + avp);
        const double ae = dt * (up - avp);
        if (m > 0 && fabs(denprev) < fabs(aw)) piv++;
        const double den = (m > 0 ? ap - aw * cprev : ap) - (m < len - 1 ? ae * e_lo : 0.0);
        const double num = (m > 0 ? aq - aw * dprev : aq) - (m < len - 1 ? ae * f_lo : 0.0);
        xm = num / den;
    }
    xm = __shfl(xm, lane & 31, 64);
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
    if (top && act) finish(node(m + 1), xm);
    // substitution outwards: upper rows m-1 .. 0, lower rows m+1 .. len-1
    double x = xm;
    const int nb = top ? m : len - 1 - m;
    for (int g0 = 0; g0 < nb; g0 += kPF2) {
        double dd[kPF2], cc[kPF2], qq[kPF2], dv[kPF2];
#pragma unroll
        for (int u = 0; u < kPF2; u++) {
            const int r = g0 + u;
            const int srow = top ? m - 1 - r : m + 1 + r;
            const bool ok = act && r < nb;
            const long p = node(srow + 1);
            dd[u] = ok ? rh[p] : 0.0;
            cc[u] = ok ? cp[p] : 0.0;
            if (D == 2) {
                qq[u] = ok ? F.q[4][L.off + p] : 0.0;
                dv[u] = ok ? F.dtev[L.off + p] : 0.0;
            }
        }
#pragma unroll
        for (int u = 0; u < kPF2; u++) {
            const int r = g0 + u;
            if (r < nb && act) {
                const int srow = top ? m - 1 - r : m + 1 + r;
                x = dd[u] - cc[u] * x;
                const long p = node(srow + 1);
                if (D < 2) {
                    rh[p] = x;
                } else {
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
        case 6: q5[hi] = q5[base + sD]; break;
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
    const dim3 gq((unsigned)((L.ni - 2 + kQX - 2) / (kQX - 1)), (unsigned)((L.nj - 2 + kQY - 2) / (kQY - 1)),
                  (unsigned)((L.nk - 2 + kc - 1) / kc));
    hipLaunchKernelGGL(k_sa_rhs, gq, dim3(kQX * kQY), 0, s, L, F, C, kc);
}

void launch_des_eddy(hipStream_t s, const Level& L, const Fields& F, const Consts& C, int itk, int kc) {
    const int b = 256;
    const unsigned gn = (unsigned)((L.n + b - 1) / b);
    launch_sa_prep(s, L, F);
    const int bl = 128;
    const unsigned gi = (unsigned)(((long)(L.nj - 2) * (L.nk - 2) + kSL - 1) / kSL);
    const unsigned gj = (unsigned)(((long)(L.ni - 2) * (L.nk - 2) + bl - 1) / bl);
    const unsigned gk = (unsigned)(((long)(L.ni - 2) * (L.nj - 2) + bl - 1) / bl);
    const unsigned gj2 = (unsigned)(((long)(L.ni - 2) * (L.nk - 2) + 63) / 64);
    const unsigned gk2 = (unsigned)(((long)(L.ni - 2) * (L.nj - 2) + 63) / 64);
    const unsigned fi = (unsigned)(((long)L.nj * L.nk + b - 1) / b);
    const unsigned fj = (unsigned)(((long)L.ni * L.nk + b - 1) / b);
    const unsigned fk = (unsigned)(((long)L.ni * L.nj + b - 1) / b);
    for (int it = 0; it < itk; it++) {
        launch_sa_rhs(s, L, F, C, kc);
        hipLaunchKernelGGL(k_sa_adi_i, dim3(gi), dim3(kSNT), 0, s, L, F);
        if (L.nj - 2 >= 4) hipLaunchKernelGGL(k_sa_adi2<1>, dim3(gj2), dim3(128), 0, s, L, F, C);
        else hipLaunchKernelGGL(k_sa_adi<1>, dim3(gj), dim3(bl), 0, s, L, F, C);
        if (L.nk - 2 >= 4) hipLaunchKernelGGL(k_sa_adi2<2>, dim3(gk2), dim3(128), 0, s, L, F, C);
        else hipLaunchKernelGGL(k_sa_adi<2>, dim3(gk), dim3(bl), 0, s, L, F, C);
        hipLaunchKernelGGL(k_sa_bc<0>, dim3(fi), dim3(b), 0, s, L, F, C);
        hipLaunchKernelGGL(k_sa_bc<1>, dim3(fj), dim3(b), 0, s, L, F, C);
        hipLaunchKernelGGL(k_sa_bc<2>, dim3(fk), dim3(b), 0, s, L, F, C);
        hipLaunchKernelGGL(k_sa_clamp, dim3(gn), dim3(b), 0, s, L, F, C);
    }
}

}  // namespace desalm
