// desalm_internal.h -- device data layout and kernel launchers of libdesalm.
//
// Layout in HBM: every field is structure-of-arrays, one double array per
// component, with all multigrid levels concatenated in the reference's node
// order (level n starts at node offset `off`, nodes run i fastest, then j,
// then k).  A kernel on level n offsets each component pointer by `off`.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/desalm.h"
#include "host_error.h"

namespace desalm {

constexpr int kMaxLevels = DESALM_MAX_LEVELS;

// Workgroups are handed to the 8 XCDs round robin (workgroup b runs on XCD
// b % 8), and every XCD has its own L2.  xcd_block(b, nb) renumbers the nb
// workgroups of a launch so that XCD x runs the x-th contiguous eighth of the
// logical grid, in order: a stencil kernel then finds the neighbouring rows
// and planes of its nodes in its own L2 instead of fetching them on every
// XCD.  A bijection of [0, nb).
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nb) {
    const unsigned x = b % 8u, l = b / 8u, q = nb / 8u, r = nb % 8u;
    return (x < r ? x * (q + 1u) : r * (q + 1u) + (x - r) * q) + l;
}

// Element idx of an array addressed as its base plus a 32-bit byte offset:
// the form of the hardware's scalar-base + vector-offset loads and stores,
// which lets the loads of several arrays at one node share one offset
// register instead of holding a 64-bit address each.  Every device array is
// below 4 GiB (desalm_init refuses more than 2^29 nodes over all levels).
template <typename T>
__device__ __forceinline__ T ld32(const T* a, unsigned idx) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(a) + idx * (unsigned)sizeof(T));
}
template <typename T>
__device__ __forceinline__ void st32(T* a, unsigned idx, T v) {
    *reinterpret_cast<T*>(reinterpret_cast<char*>(a) + idx * (unsigned)sizeof(T)) = v;
}

struct Level {
    int ni, nj, nk;
    long off;   // node offset of the level in all-level arrays
    long n;     // ni * nj * nk
    double dc, de, dz;
};

// Run constants used inside kernels (a trimmed copy of desalm_params plus
// the derived sa/sb tables of init_daf_bc).
struct Consts {
    double ren, rei, beta, delti, e_source;
    double cfl1, vnn1, cfl2, vnn2, pdiss;
    double ep[2][3];
    double sa[6][5], sb[6][5];
    int btype[6];
    int unsteady;
    int mixed;     // DESALM_PREC_MIXED: DAF sweeps in FP32 on the *f scratch
    int dtau_met;  // DESALM_DTAU_MET: viscous dtau limits without nu_t
    double fb[3];  // body force per unit mass (x, y, z), -fb / J in rh(2:4)
};

// Device pointers.  Fields marked "all" span every level; "fine" span the
// fine level only (scratch is only ever used one level at a time and the
// fine level is the largest).
struct Fields {
    double* csi[3];   // all
    double* eta[3];   // all
    double* zet[3];   // all
    double* aj;       // all
    double* q[5];     // all
    double* qn[5];    // all
    double* qnm1[5];  // all
    double* qold[5];  // all (qold_mg)
    double* rh[4];    // all
    double* pk[4];    // all
    double* xnut;     // all
    double* wd;       // all (zero off the fine level)
    double* dtev;     // all (zero off the fine level)
    double* x[3];     // all: grid coordinates
    // scratch
    double* dtau;     // all: local pseudo time step of the level being smoothed
    double* diss1;    // all: pressure dissipation kept for the recalc branch
    double* cp[4];    // fine: line-solve c' per scalar system
    double* dp[4];    // fine: sweep right-hand sides / d', prolongation dq
    double* av[3];    // fine: SA ADI face coefficients per direction
    double* dtj;      // fine: SA ADI 0.5 dtev J
    double* ucn[3];   // fine: contravariant velocity / J (residual, SA update)
    double* pj;       // fine: p / J
    double* fv[3];    // fine: pressure-dissipation fv per direction
    // FP32 DAF scratch of the mixed-precision form (fine size; null in FP64)
    float* dpf[4];
    float* cpf[4];
    double* wo;       // fine: SA vorticity measure
    double* rh5;      // fine: SA right-hand side
    double* resid_part;  // [kResidBlocks * 5] partial L1 sums
    double* resid;       // [5] L1 sums
    double* pfix;        // [1] pinned pressure value
    // inflow (btype 4) and characteristic-outlet (btype 5) faces of the i and
    // j directions: grid ratio of the extrapolation per face node (null where
    // the face is neither; the k faces use 0, see k_bc_fm)
    double* rat[4];
    unsigned long long* pivots;  // [1] rows where LAPACK SGTSV would pivot
    unsigned* zero_pivot;        // [1] first line row with a zero or non-finite pivot (0xffffffff: none)
};

constexpr int kResidBlocks = 4096;

// Overset donor table of the fine level (device pointers, level-local node
// indices): donor d reads the cell whose lower corner is host[d] with
// trilinear weights coef[3d..3d+2] into buf; fringe node node[d] takes the
// value in buffer slot slot[d] (slot d when slot is null).
struct Overset {
    long n;
    const long* host;
    const double* coef;
    const long* node;
    double* buf;        // 4 values per slot: gathered donor values
    const long* slot;
};

// Actuator-line rotor on the device (desalm_turbine of desalm.h)
struct AlmTurbine {
    int nblades, nelems, nangles, itur1, itur2, n_disk, n_ker;
    double tur_rad, hub_rad, tsr, angle, dtheta;
    const int* node_idx;     // [nelems * nangles * 3]
    const double* node_pos;  // [nelems * nangles * 27]
    const double* disk_dat;  // [jmg * kmg * 8]
    const int* pol_b;        // [jmg * kmg]
    const int* pol_a;
    const int* disk_nodes;   // [n_disk * 3]
    const int* ker_nodes;    // [n_ker * 3]
    const double* ker_pos;
    const double* ker_dat;
    double* forces;          // [nblades * nelems * 6] position and force per element
};

// all polars, concatenated: polar p is points off[p] .. off[p] + cnt[p] - 1
struct AlmPolars {
    const int* off;
    const int* cnt;
    const double* angle;
    const double* cl;
    const double* cd;
};

struct Alm {
    std::vector<AlmTurbine> t;
    AlmPolars pol{nullptr, nullptr, nullptr, nullptr, nullptr};
    double length_scale = 0.0, vel_infty = 0.0;
    int jmg = 0, kmg = 0;
    std::vector<void*> dev;  // device allocations of the tables
};

// Learned warm start (k_mlp.hip): per-node features are the last two time
// differences of the five variables at the node and its six face
// neighbours (each variable's 14 divided by their root mean square), plus
// q^n at the node; the network's 10 outputs are, per variable, the
// coefficients' departures (a - 1, b) of the start
// q^n + a (q^n - q^(n-1)) + b (q^(n-1) - q^(n-2)).
constexpr int kWsComp = 5;
constexpr int kWsPts = 9;                              // the node and 8 neighbours
constexpr int kWsIn = kWsComp * 2 * kWsPts + kWsComp;  // 95
constexpr int kWsOut = 2 * kWsComp;                     // network outputs
constexpr int kWsTgt = 3 * kWsComp;                     // per sample: r, d1, d2
constexpr int kWsStats = 2 * kWsIn + 2 * kWsComp;       // mu, 1/sd, scale, unused
constexpr int kWsMaxHidden = 64;
struct WsQ {
    double* q[5];  // q^(n-2), fine level
};

// Random-flow-generator inflow (k_rfg.hip): N random Fourier modes and, per
// (j, k) of the inflow plane (index j + nj k), the mean velocity, the
// eigenvectors (column-major 3 x 3) and the square roots of the eigenvalues
// of the Reynolds-stress tensor, and the time and length scales
struct Rfg {
    int n;
    const double* k;      // [3 n]
    const double* omega;  // [n]
    const double* p;      // [3 n]
    const double* q;      // [3 n]
    const double* um;     // [3 nj nk]
    const double* m;      // [9 nj nk]
    const double* c;      // [3 nj nk]
    const double* ts;     // [nj nk]
    const double* ls;     // [nj nk]
};

struct Ctx;  // host context, desalm_api.hip

// ---- one launch over the zones of a group ----
// The block-descriptor table: every zone's levels, fields and constants in
// device memory.  A batched launch covers the workgroups of all zones, zone
// z owning workgroups start[z] .. start[z+1]-1; a workgroup finds its zone by
// a scan of the (at most kMaxBatchZones + 1) starts, all in its kernel
// arguments, and runs the single-zone body with that zone's level, fields
// and constants and its zone-local workgroup index.
constexpr int kMaxBatchZones = 16;
struct ZoneDesc {
    Level lv[kMaxLevels + 1];
    Fields F;
    Consts C;
    long pfix;  // global node of the pinned pressure (bcond_fm.F90:310-322)
    int icrs, jcrs, kcrs;
};
struct ZoneGrid {
    const ZoneDesc* __restrict__ z;
    int nz;
    unsigned start[kMaxBatchZones + 1];
};
struct ZoneBlock {
    int z;
    unsigned bid, nblk;
};
__device__ __forceinline__ ZoneBlock zone_block(const ZoneGrid& G) {
    const unsigned b = blockIdx.x;
    int z = 0;
    while (z + 1 < G.nz && b >= G.start[z + 1]) z++;
    return {z, b - G.start[z], G.start[z + 1] - G.start[z]};
}
// host side of a table: the host copies (block counts are computed from
// them) and the device copy
struct ZoneList {
    std::vector<ZoneDesc> host;
    ZoneDesc* dev = nullptr;
    int nz() const { return (int)host.size(); }
    // grid of one launch; nb(zone) = that zone's workgroup count
    template <class NB>
    ZoneGrid grid(NB nb, unsigned& total) const {
        ZoneGrid G{};
        G.z = dev;
        G.nz = nz();
        unsigned acc = 0;
        for (int z = 0; z < G.nz; z++) {
            G.start[z] = acc;
            acc += (unsigned)nb(host[z]);
        }
        G.start[G.nz] = acc;
        total = acc;
        return G;
    }
};

// ---- launchers (each enqueues on `s`) ----
// residual assembly, k_residual.hip
// dtau and fv of level n; on a fine level with a characteristic outlet also
// p/J and U/J for it, U/J unless ucn_kept (already in F.ucn, stored by the SA
// update of the same pseudo-iteration)
void launch_dtau(hipStream_t s, const Level& L, const Fields& F, const Consts& C, int n, bool ucn_kept);
// kc: k planes per workgroup of the LDS-tiled kernels (0: default).  The
// recalc branch reads p as it is: a bcond_fm before it applies the
// pinned-pressure shift itself (pfix_deferred false).
void launch_residual(hipStream_t s, const Level& L, const Fields& F, const Consts& C, int n,
                     int recalc, int calc_pk, int kc);
void launch_visc_pk(hipStream_t s, const Level& L, const Fields& F, const Consts& C, int n, int calc_pk, int kc);
int visc_kc_default(const Level& L);
// DAF implicit operator, k_daf.hip
void launch_daf_i(hipStream_t s, const Level& L, const Fields& F, const Consts& C, int n);
void launch_daf_j(hipStream_t s, const Level& L, const Fields& F, const Consts& C, int n);
void launch_daf_k(hipStream_t s, const Level& L, const Fields& F, const Consts& C, int n,
                  double* dq_out /* nullable: also store Mc x here (4 SoA, level-local) */);
// fine-grid BCs, k_mg.hip
void launch_outlet_ratios(hipStream_t s, const Level& L, const Fields& F, const int* btype, const int* bdir);
// right-hand side of the characteristic-outlet planes (btype 5 on i faces,
// fine level): mg_brhs + nonreflect_ibc (k_outlet.hip)
void launch_outlet_rhs(hipStream_t s, const Level& L, const Fields& F, const Consts& C);
// pfix_deferred: read the pinned pressure into F.pfix but leave its
// subtraction to the next pass over the level (launch_resid_and_save with
// pfix), which then reads p first
void launch_bcond_fm(hipStream_t s, const Level& L, const Fields& F, const Consts& C, int ifix,
                     int jfix, int kfix, bool pfix_deferred = false);
// SA / DES, k_sa.hip
void launch_des_eddy(hipStream_t s, const Level& L, const Fields& F, const Consts& C, int itk, int kc);
void launch_sa_prep(hipStream_t s, const Level& L, const Fields& F);
void launch_sa_rhs(hipStream_t s, const Level& L, const Fields& F, const Consts& C, int kc);
int sa_kc_default(const Level& L);
// multigrid, k_mg.hip
void launch_inject(hipStream_t s, const Level* lv, int ng, int icrs, int jcrs, int kcrs,
                   double* const* comps, int ncomp);
void launch_nlevel(hipStream_t s, const Level& Lc, const Level& Lf, const Fields& F, int icrs,
                   int jcrs, int kcrs);
void launch_zero_pk_boundary(hipStream_t s, const Level& L, const Fields& F);
void launch_prolong(hipStream_t s, const Level& Lf, const Level& Lc, const Fields& F,
                    const Consts& C, int icrs, int jcrs, int kcrs);
void launch_resid_and_save(hipStream_t s, const Level& L, const Fields& F, int save_only, bool pfix = false);
void launch_copy_qold(hipStream_t s, const Level& L, const Fields& F);
// physical-step outputs: the L2 sums of q - qn over the fine interior into
// F.resid (src/main.F90:214-300), and q(1:5), xnut at fine-level nodes
void launch_real_residual(hipStream_t s, const Level& L, const Fields& F);
void launch_probe(hipStream_t s, const Fields& F, int n, const long* node, double* out);
// case hooks of mg_driver (k_hooks.hip): the outlet mass-flux rescale and
// the wall functions of both j walls, on the fine level
void launch_conservation(hipStream_t s, const Level& L, const Fields& F);
void launch_wall_functions(hipStream_t s, const Level& L, const Fields& F, double ren);
// metrics, k_metrics.hip
void launch_metrics(hipStream_t s, const Level& L, const Fields& F, const Consts& C, int fine);
// layout conversion, k_mg.hip
void launch_aos_to_soa(hipStream_t s, const double* aos, int ncomp, long n, double* const* soa);
void launch_soa_to_aos(hipStream_t s, double* const* soa, int ncomp, long n, double* aos);
// overset interpolation (tur = 0: q(1:4), 1: nu~ and xnut), k_overset.hip
void launch_bintp(hipStream_t s, const Level& L, const Fields& F, const Consts& C, const Overset& O, int tur,
                  double cvin, int des);
void launch_bintp_gather(hipStream_t s, const Level& L, const Fields& F, const Overset& O, int tur);
void launch_bintp_scatter(hipStream_t s, const Level& L, const Fields& F, const Consts& C, const Overset& O, int tur,
                          double cvin, int des);
// actuator-line model, k_alm.hip: forcing added to rh of level L, and
// the hub zeroing of bcond_fm on the fine level
void launch_alm_rhs(hipStream_t s, const Level& L, const Fields& F, const Alm& A, double ren);
void launch_rfg_inflow(hipStream_t s, const Level& L, const Fields& F, const Rfg& R, double t);
void launch_alm_hub(hipStream_t s, const Level& L, const Fields& F, const Alm& A);
// every zone of a group in one launch per kernel (the table's level n)
void launch_dtau_z(hipStream_t s, const ZoneList& Z, int n, bool ucn_kept);
void launch_residual_z(hipStream_t s, const ZoneList& Z, int n, int recalc, int calc_pk, int kc);
void launch_daf_z(hipStream_t s, const ZoneList& Z, int n, int D);
void launch_des_eddy_z(hipStream_t s, const ZoneList& Z, int itk, int kc);
void launch_nlevel_z(hipStream_t s, const ZoneList& Z, int n);
void launch_prolong_z(hipStream_t s, const ZoneList& Z, int n);
void launch_bcond_fm_z(hipStream_t s, const ZoneList& Z);
void launch_resid_and_save_z(hipStream_t s, const ZoneList& Z, int save);
// batched GEPP (LAPACK SGTSV semantics), k_daf.hip
void launch_gtsv_batch(hipStream_t s, int n, int nsys, double* dl, double* d, double* du,
                       double* b, int* info);

void launch_ws_sample(hipStream_t s, const Level& L, const Fields& F, const WsQ& Q, float* X, float* T, long m,
                      unsigned long long seed, long step);
void launch_ws_batch(hipStream_t s, const float* X, const float* T, long n, float* Xb, float* Tb, int m,
                     unsigned long long seed, long it, const float* stats);
void launch_ws_features(hipStream_t s, const Level& L, const Fields& F, const WsQ& Q, float* X, long r0, long m,
                        const float* stats);
void launch_ws_apply(hipStream_t s, const Level& L, const Fields& F, const WsQ& Q, const float* Y, long r0, long m);
void launch_dense(hipStream_t s, const float* A, int lda, int K1, const float* W, int N, float* Y, int ldy, int M,
                  int act, int ones);
void launch_dense_bwd(hipStream_t s, const float* dZ, int ldz, int N, const float* W, int K, const float* Aprev,
                      int lda, float* dA, int ldd, int M);
int wgrad_parts(int M);
void launch_wgrad(hipStream_t s, const float* A, int lda, int K1, const float* dZ, int ldz, int N, int M, float* P,
                  float* G);
void launch_ws_loss(hipStream_t s, const float* Y, const float* T, int M, float* dZ, double* part, int nparts);
void launch_adam(hipStream_t s, float* w, const float* g, float* m, float* v, int n, float lr, float b1, float b2,
                 float eps, float c1, float c2);
void launch_colstats(hipStream_t s, const float* X, int ld, int nc, long n, double* part, int nparts);

}  // namespace desalm
