// desalm_api.hip -- C ABI of libdesalm.so: context, device memory, the
// drop-in entry points for mg_driver / solver_daf / des_eddy and the
// V-cycle orchestration (src/mg_driver.F90:2-323).
//
// All state of a context lives on its device for the context's lifetime;
// host arrays are only borrowed for the duration of a call.  Every kernel is
// enqueued on one stream; the only host synchronisation inside
// desalm_mg_driver is the 5-double read-back of the update norms per
// pseudo-iteration that the convergence test needs.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "desalm_internal.h"

using namespace desalm;

// Per-kernel device timing with HIP events on the context stream.
enum KernelId {
    KID_DTAU, KID_RESIDUAL, KID_RESIDUAL_RECALC, KID_DAF_I, KID_DAF_J, KID_DAF_K, KID_BCOND,
    KID_DES_EDDY, KID_NLEVEL, KID_PROLONG, KID_RESID, KID_OVERSET, KID_ALM, KID_COUNT
};
static const char* kKernelNames[KID_COUNT] = {
    "dtau", "residual", "residual_recalc", "daf_i", "daf_j", "daf_k", "bcond_fm",
    "des_eddy", "nlevel", "prolong", "resid_norm", "overset", "alm"};

struct Profiler {
    bool on = false;
    struct Rec {
        int id, level;
        hipEvent_t a, b;
    };
    std::vector<Rec> pending;
    std::vector<hipEvent_t> pool;
    double ms[KID_COUNT][kMaxLevels + 1] = {};
    long long count[KID_COUNT][kMaxLevels + 1] = {};
};

// Learned warm start state (desalm_ws_*): the sample buffer, the network's
// packed parameters with their Adam moments, the normalisation statistics
// and the scratch matrices of training and inference.  Parameter layout:
// W1 (kWsIn + 1) x nh, W2 (nh + 1) x nh, W3 (nh + 1) x kWsOut, row-major,
// the last row of each the bias.
struct WarmStart {
    int mode = DESALM_WS_NONE;
    int nh = 0;
    long per_step = 0, cap = 0, rows = 0, records = 0;
    unsigned long long seed = 0;
    WsQ q2{{nullptr, nullptr, nullptr, nullptr, nullptr}};  // q^(n-2), fine level
    float* X = nullptr;   // cap x (kWsIn + 1) raw features
    float* T = nullptr;   // cap x kWsTgt raw targets (r, d1, d2)
    float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    long adam_t = 0;
    float* stats = nullptr;  // kWsStats: feature mu, 1/sd (kWsIn each), target scale (kWsComp), unused
    bool have_stats = false;
    int bcap = 0;  // rows of the scratch matrices
    float *A0 = nullptr, *A1 = nullptr, *A2 = nullptr, *Y = nullptr, *Tb = nullptr;
    float *dZ3 = nullptr, *dZ2 = nullptr, *dZ1 = nullptr, *P = nullptr;
    double* lpart = nullptr;  // loss partial sums, kWsLossParts per training iteration
    long lcap = 0;
    std::vector<void*> dev;
    long long bytes = 0;
    int np() const { return (kWsIn + 1) * nh + (nh + 1) * nh + (nh + 1) * kWsOut; }
    int o2() const { return (kWsIn + 1) * nh; }
    int o3() const { return (kWsIn + 1) * nh + (nh + 1) * nh; }
};
constexpr int kWsLossParts = 64;

struct desalm_ctx {
    desalm_params p;
    int device = 0;
    hipStream_t stream = nullptr;
    int ng = 0;
    Level lv[kMaxLevels + 1];
    long nall = 0, n1 = 0, maxdim = 0;
    Consts C;
    Fields F;
    std::vector<void*> allocs;
    // the fields' allocation (dalloc)
    char* arena = nullptr;
    size_t arena_cap = 0, arena_used = 0;
    long arena_n = 0;
    bool sizing = false;
    double* staging = nullptr;     // 5 * nall doubles: AoS transfers
    double* dq_buf = nullptr;      // 4 * n1 doubles: dq of the solver_daf parity entry
    std::string err;
    double last_ms = 0.0;
    unsigned long long last_pivots = 0;
    double init_resd[5] = {0, 0, 0, 0, 0};
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    Profiler prof;
    // overset donor table (desalm_set_overset); its buffers are freed on
    // replacement and by desalm_finalize
    Overset ov{0, nullptr, nullptr, nullptr, nullptr, nullptr};
    int ov_mom = 0;
    double cvin = 0.0;
    Alm alm;  // actuator-line rotors (desalm_set_turbines)
    WarmStart ws;
    // k planes per workgroup of the LDS-tiled viscous kernel per level and
    // of the SA right-hand side (0: default; desalm_autotune / desalm_set_tiles)
    int kc_visc[kMaxLevels + 1] = {};
    int kc_sa = 0;
    long long dev_bytes = 0;  // fields and scratch (dalloc)
    long long alm_bytes = 0;  // rotor tables
    // random-flow-generator inflow (desalm_set_rfg); applied by desalm_step
    Rfg rfg{0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    std::vector<void*> rfg_dev;
    double rfg_time0 = 0.0;
    long long steps = 0;  // desalm_step calls since desalm_set_rfg
    // case hooks of mg_driver (DESALM_HOOK_*, desalm_set_hooks); the hooks
    // read the grid coordinates, so they need desalm_set_grid
    int hooks = 0;
    bool have_grid = false;
    // failure detection of the last driver call (desalm_last_status) and the
    // entry state a non-plain form falls back from (qnm1 5, q^(n-2) 5, xnut)
    int st_zero_pivot = 0, st_nonfinite = 0, st_diverged = 0;
    long long fallbacks = 0;
    double* fb_save = nullptr;
    // desalm_probe: node indices and values of up to probe_cap points
    void* probe_dev = nullptr;
    int probe_cap = 0;
};

namespace {

// growth of the largest u, v, w update norm over the first iteration's, in
// decades, beyond which a step counts as diverged
constexpr double kDivergeDecades = 3.0;

thread_local std::string g_err;  // errors without a context (init, gtsv)

int fail(desalm_ctx* c, const std::string& msg) {
    if (c) c->err = msg;
    g_err = msg;
    return -1;
}

#define HIPCHK(c, x)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (x);                                                                \
        if (e_ != hipSuccess)                                                               \
            return fail((c), std::string(#x " failed: ") + hipGetErrorString(e_));        \
    } while (0)

// Device fields are carved from one allocation: a k-line sweep touches a
// dozen arrays at rows about 200 KB apart, and separate allocations cost it
// an address translation per array and row (a sixth of its accesses missed
// the L1 translation cache).  Consecutive arrays start a further 4 KiB +
// 256 B into the allocation (16 steps, then again from zero), so that the
// same node of different arrays falls on different L2 channels.  With
// c->sizing set, dalloc only adds up the bytes the arena needs.
constexpr size_t kStaggerWords = 544;  // 4352 B

double* dalloc(desalm_ctx* c, long n) {
    const size_t words = (size_t)(n > 0 ? n : 1);
    const size_t pad = (size_t)(c->arena_n++ % 16) * kStaggerWords;
    const size_t bytes = ((pad + words) * sizeof(double) + 255) & ~(size_t)255;
    if (c->sizing) {
        c->arena_used += bytes;
        return nullptr;
    }
    c->dev_bytes += (long long)bytes;
    if (c->arena && c->arena_used + bytes <= c->arena_cap) {
        double* q = reinterpret_cast<double*>(c->arena + c->arena_used) + pad;
        c->arena_used += bytes;
        return q;
    }
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
    c->allocs.push_back(p);
    if (hipMemset(p, 0, bytes) != hipSuccess) return nullptr;
    return (double*)p + pad;
}

int validate(const desalm_params* p, std::string& why) {
    if (p->ng < 1 || p->ng > kMaxLevels) { why = "ng must be in 1..6"; return -1; }
    if ((p->icrs | p->jcrs | p->kcrs) & ~1) { why = "icrs/jcrs/kcrs must be 0 or 1"; return -1; }
    int i = p->img, j = p->jmg, k = p->kmg;
    for (int n = 1; n <= p->ng; n++) {
        if (i < 5 || j < 5 || k < 5) { why = "every level needs at least 5 nodes per direction"; return -1; }
        if (n < p->ng) {
            if ((p->icrs && !(i & 1)) || (p->jcrs && !(j & 1)) || (p->kcrs && !(k & 1))) {
                why = "coarsened directions need an odd node count on every level but the coarsest";
                return -1;
            }
            i = (i + p->icrs) / (1 << p->icrs);
            j = (j + p->jcrs) / (1 << p->jcrs);
            k = (k + p->kcrs) / (1 << p->kcrs);
        }
    }
    if (p->img > 4098 || p->jmg > 4098 || p->kmg > 4098) {
        why = "at most 4098 nodes per direction (lines of up to 4096 rows)";
        return -1;
    }
    {
        // kernels address every array with 32-bit byte offsets
        long nall = 0, a = p->img, b = p->jmg, d = p->kmg;
        for (int n = 1; n <= p->ng; n++) {
            nall += a * b * d;
            a = (a + p->icrs) / (1 << p->icrs);
            b = (b + p->jcrs) / (1 << p->jcrs);
            d = (d + p->kcrs) / (1 << p->kcrs);
        }
        if (nall > (1L << 29)) {
            why = "at most 2^29 nodes over all levels (arrays below 4 GiB)";
            return -1;
        }
    }
    for (int b = 0; b < 6; b++)
        if (p->btype[b] < 0 || p->btype[b] > 8) {
            why = "unsupported boundary type (btype 0-8)";
            return -1;
        }
    for (int d = 0; d < 3; d++)
        if ((p->btype[2 * d] == DESALM_BC_PERIODIC_WRAP) != (p->btype[2 * d + 1] == DESALM_BC_PERIODIC_WRAP)) {
            why = "a periodic-wrap direction (btype 7) needs btype 7 on both of its faces";
            return -1;
        }
    for (int c = 0; c < 3; c++)
        if (!std::isfinite(p->body_force[c])) {
            why = "body force must be finite";
            return -1;
        }
    for (int b = 0; b < 4; b++)
        if ((p->btype[b] == DESALM_BC_INFLOW || p->btype[b] == DESALM_BC_OUTLET_MOC) &&
            (p->bdir[b] < 1 || p->bdir[b] > 3)) {
            why = "an inflow or characteristic-outlet face (btype 4, 5) needs bdir 1, 2 or 3: its extrapolation "
                  "uses that coordinate";
            return -1;
        }
    if (p->ifix < 1 || p->ifix > p->img || p->jfix < 1 || p->jfix > p->jmg || p->kfix < 1 || p->kfix > p->kmg) {
        why = "pinned-pressure node outside the fine level";
        return -1;
    }
    if (p->itk < 0 || p->itm < 0) { why = "itk and itm must be non-negative"; return -1; }
    if (p->precision != DESALM_PREC_FP64 && p->precision != DESALM_PREC_MIXED) {
        why = "precision must be DESALM_PREC_FP64 or DESALM_PREC_MIXED";
        return -1;
    }
    return 0;
}

hipEvent_t prof_event(desalm_ctx* c) {
    if (!c->prof.pool.empty()) {
        hipEvent_t e = c->prof.pool.back();
        c->prof.pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Runs `launch` between two events when profiling is on.
template <class F>
void timed(desalm_ctx* c, int id, int level, F&& launch) {
    if (!c->prof.on) {
        launch();
        return;
    }
    hipEvent_t a = prof_event(c), b = prof_event(c);
    if (!a || !b) {
        launch();
        return;
    }
    (void)hipEventRecord(a, c->stream);
    launch();
    (void)hipEventRecord(b, c->stream);
    c->prof.pending.push_back({id, level, a, b});
}

// Accumulates finished records; call after the stream has been synchronised.
void prof_flush(desalm_ctx* c) {
    for (auto& r : c->prof.pending) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            c->prof.ms[r.id][r.level] += ms;
            c->prof.count[r.id][r.level] += 1;
        }
        c->prof.pool.push_back(r.a);
        c->prof.pool.push_back(r.b);
    }
    c->prof.pending.clear();
}

// bcond_fm on the fine level (src/bcond_fm.F90:2-362), with the
// actuator-line hub zeroing at its end (:328-355); pfix_deferred leaves the
// pinned-pressure shift to the caller's next pass over the level
void bcond(desalm_ctx* c, hipStream_t s, bool pfix_deferred = false) {
    launch_bcond_fm(s, c->lv[1], c->F, c->C, c->p.ifix, c->p.jfix, c->p.kfix, pfix_deferred);
    if (!c->alm.t.empty()) launch_alm_hub(s, c->lv[1], c->F, c->alm);
}

// ucn_kept: des_eddy has just stored U/J of the current velocities (fine
// level, SA/DES on; it changes only nu~ and xnut)
void solver_daf(desalm_ctx* c, int n, int recalc, int calc_pk, double* dq_out, bool ucn_kept = false) {
    const Level& L = c->lv[n];
    hipStream_t s = c->stream;
    timed(c, KID_DTAU, n, [&] { launch_dtau(s, L, c->F, c->C, n, ucn_kept); });
    timed(c, KID_RESIDUAL, n, [&] { launch_residual(s, L, c->F, c->C, n, 0, calc_pk, c->kc_visc[n]); });
    // characteristic-outlet planes of the fine level (solver_daf.F90:322-365)
    if (n == 1 && (c->C.btype[0] == DESALM_BC_OUTLET_MOC || c->C.btype[1] == DESALM_BC_OUTLET_MOC))
        timed(c, KID_RESIDUAL, n, [&] { launch_outlet_rhs(s, L, c->F, c->C); });
    // actuator-line forcing on every level (solver_daf.F90:378-381)
    if (!c->alm.t.empty()) timed(c, KID_ALM, n, [&] { launch_alm_rhs(s, L, c->F, c->alm, c->C.ren); });
    timed(c, KID_DAF_I, n, [&] { launch_daf_i(s, L, c->F, c->C, n); });
    timed(c, KID_DAF_J, n, [&] { launch_daf_j(s, L, c->F, c->C, n); });
    timed(c, KID_DAF_K, n, [&] { launch_daf_k(s, L, c->F, c->C, n, dq_out); });
    if (n == 1) timed(c, KID_BCOND, n, [&] { bcond(c, s); });
    if (recalc == 1)
        timed(c, KID_RESIDUAL_RECALC, n, [&] { launch_residual(s, L, c->F, c->C, n, 1, 0, c->kc_visc[n]); });
}

int copy_in_level(desalm_ctx* c, double* const* soa, int ncomp, const double* host, const Level& L) {
    if (!host) return 0;
    HIPCHK(c, hipMemcpyAsync(c->staging, host, sizeof(double) * ncomp * L.n, hipMemcpyHostToDevice, c->stream));
    double* dst[8];
    for (int m = 0; m < ncomp; m++) dst[m] = soa[m] + L.off;
    launch_aos_to_soa(c->stream, c->staging, ncomp, L.n, dst);
    HIPCHK(c, hipGetLastError());
    return 0;
}

int copy_out_level(desalm_ctx* c, double* const* soa, int ncomp, double* host, const Level& L) {
    if (!host) return 0;
    double* src[8];
    for (int m = 0; m < ncomp; m++) src[m] = soa[m] + L.off;
    launch_soa_to_aos(c->stream, src, ncomp, L.n, c->staging);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(host, c->staging, sizeof(double) * ncomp * L.n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

}  // namespace

namespace desalm {
int host_fail(const std::string& msg) { return fail(nullptr, msg); }
}  // namespace desalm

extern "C" {

int desalm_init(const desalm_params* p, int device, desalm_ctx** out) {
    if (!p || !out) return fail(nullptr, "null argument");
    *out = nullptr;
    std::string why;
    if (validate(p, why)) return fail(nullptr, why);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0)
        return fail(nullptr, "no HIP device " + std::to_string(device));
    desalm_ctx* c = new desalm_ctx();
    c->p = *p;
    c->device = device;
    c->ng = p->ng;
    if (hipSetDevice(device) != hipSuccess) { delete c; return fail(nullptr, "hipSetDevice failed"); }
    if (hipStreamCreate(&c->stream) != hipSuccess) { delete c; return fail(nullptr, "hipStreamCreate failed"); }
    if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        desalm_finalize(c);
        return fail(nullptr, "hipEventCreate failed");
    }
    // levels (src/init.F90:254-260, init_mg :1116-1120)
    int i = p->img, j = p->jmg, k = p->kmg;
    long off = 0;
    for (int n = 1; n <= c->ng; n++) {
        Level& L = c->lv[n];
        L.ni = i; L.nj = j; L.nk = k;
        L.off = off;
        L.n = (long)i * j * k;
        L.dc = 1.0 / pow(2.0, (double)((n - 1) * p->icrs));
        L.de = 1.0 / pow(2.0, (double)((n - 1) * p->jcrs));
        L.dz = 1.0 / pow(2.0, (double)((n - 1) * p->kcrs));
        off += L.n;
        long md = i > j ? i : j;
        md = md > k ? md : k;
        if (md > c->maxdim) c->maxdim = md;
        i = (i + p->icrs) / (1 << p->icrs);
        j = (j + p->jcrs) / (1 << p->jcrs);
        k = (k + p->kcrs) / (1 << p->kcrs);
    }
    c->nall = off;
    c->n1 = c->lv[1].n;
    // constants (init_input_file, init_daf_bc)
    Consts& C = c->C;
    memset(&C, 0, sizeof(C));
    C.ren = p->ren;
    C.rei = 1.0 / p->ren;
    c->cvin = 0.1 / p->ren;  // cvin = pt1 / ren (src/init.F90:250)
    C.beta = p->beta;
    C.delti = p->delti;
    C.e_source = p->unsteady ? 1.0 : 0.0;
    C.unsteady = p->unsteady ? 1 : 0;
    C.mixed = p->precision == DESALM_PREC_MIXED ? 1 : 0;
    C.dtau_met = p->dtau_formula == DESALM_DTAU_MET ? 1 : 0;
    for (int m = 0; m < 3; m++) C.fb[m] = p->body_force[m];
    C.cfl1 = p->cfl1; C.vnn1 = p->vnn1; C.cfl2 = p->cfl2; C.vnn2 = p->vnn2;
    C.pdiss = p->pdiss_coef;
    for (int a = 0; a < 2; a++)
        for (int d = 0; d < 3; d++) C.ep[a][d] = p->ep[a][d];
    for (int b = 0; b < 6; b++) {
        C.btype[b] = p->btype[b];
        for (int m = 0; m < 5; m++) {
            C.sa[b][m] = 4.0 / 3.0;
            C.sb[b][m] = -(1.0 / 3.0);
        }
        if (p->btype[b] == 1)
            for (int m = 1; m < 4; m++) C.sa[b][m] = C.sb[b][m] = 0.0;
        if (p->btype[b] == 2 && p->bdir[b] >= 1 && p->bdir[b] <= 3) C.sa[b][p->bdir[b]] = C.sb[b][p->bdir[b]] = 0.0;
    }
    // device fields
    Fields& F = c->F;
    memset(&F, 0, sizeof(F));
    bool ok = true;
    auto A = [&](long n) {
        double* q = dalloc(c, n);
        if (!q && !c->sizing) ok = false;
        return q;
    };
    auto fields = [&] {
        for (int m = 0; m < 3; m++) { F.csi[m] = A(c->nall); F.eta[m] = A(c->nall); F.zet[m] = A(c->nall); F.x[m] = A(c->nall); }
        F.aj = A(c->nall);
        for (int m = 0; m < 5; m++) { F.q[m] = A(c->nall); F.qn[m] = A(c->nall); F.qnm1[m] = A(c->nall); F.qold[m] = A(c->nall); }
        for (int m = 0; m < 4; m++) { F.rh[m] = A(c->nall); F.pk[m] = A(c->nall); }
        F.xnut = A(c->nall); F.wd = A(c->nall); F.dtev = A(c->nall);
        F.dtau = A(c->nall); F.diss1 = A(c->nall);
        for (int m = 0; m < 4; m++) F.cp[m] = A(c->n1);
        for (int m = 0; m < 4; m++) F.dp[m] = A(c->n1);
        for (int m = 0; m < 3; m++) F.av[m] = A(c->n1);
        F.dtj = A(c->n1);
        for (int m = 0; m < 3; m++) F.ucn[m] = A(c->n1);
        F.wo = A(c->n1); F.rh5 = A(c->n1); F.pj = A(c->n1);
        for (int m = 0; m < 3; m++) F.fv[m] = A(c->n1);
        if (c->C.mixed) {
            // FP32 DAF scratch: 13 floats per fine node, carved from doubles
            auto Af = [&](long n) { return (float*)A((n + 1) / 2); };
            for (int m = 0; m < 4; m++) { F.dpf[m] = Af(c->n1); F.cpf[m] = Af(c->n1); }
        }
        F.resid_part = A(5L * kResidBlocks);
        // the five L1 sums, then the first zero or non-finite pivot of the line
        // solves (one word, read back with the sums)
        F.resid = A(6);
        F.pfix = A(1);
        for (int b = 0; b < 4; b++) {
            const Level& L1 = c->lv[1];
            F.rat[b] = (p->btype[b] == DESALM_BC_INFLOW || p->btype[b] == DESALM_BC_OUTLET_MOC)
                           ? A(b < 2 ? (long)L1.nj * L1.nk : (long)L1.ni * L1.nk)
                           : nullptr;
        }
        c->staging = A(5 * c->nall);
        c->dq_buf = A(4 * c->n1);
    };
    // first the arena's size, then the arrays carved from it
    c->sizing = true;
    fields();
    c->sizing = false;
    {
        void* pa = nullptr;
        if (hipMalloc(&pa, c->arena_used) != hipSuccess || hipMemset(pa, 0, c->arena_used) != hipSuccess) {
            if (pa) (void)hipFree(pa);
            desalm_finalize(c);
            return fail(nullptr, "device allocation failed");
        }
        c->allocs.push_back(pa);
        c->arena = (char*)pa;
        c->arena_cap = c->arena_used;
        c->arena_used = 0;
        c->arena_n = 0;
    }
    fields();
    if (ok) F.zero_pivot = (unsigned*)(F.resid + 5);
    {
        void* pv = nullptr;
        if (hipMalloc(&pv, sizeof(unsigned long long)) != hipSuccess) ok = false;
        else {
            c->allocs.push_back(pv);
            F.pivots = (unsigned long long*)pv;
            if (hipMemset(pv, 0, sizeof(unsigned long long)) != hipSuccess) ok = false;
        }
    }
    if (!ok) {
        desalm_finalize(c);
        return fail(nullptr, "device allocation failed");
    }
    if (hipDeviceSynchronize() != hipSuccess) {
        desalm_finalize(c);
        return fail(nullptr, "device synchronisation failed after allocation");
    }
    *out = c;
    return 0;
}

int64_t desalm_nodes_all(const desalm_ctx* c) { return c ? c->nall : -1; }

int64_t desalm_level_offset(const desalm_ctx* c, int n) {
    if (!c || n < 1 || n > c->ng) return -1;
    return c->lv[n].off;
}

int desalm_set_grid(desalm_ctx* c, const double* x, const double* y, const double* z) {
    if (!c || !x || !y || !z) return fail(c, "null argument");
    HIPCHK(c, hipSetDevice(c->device));
    const double* h[3] = {x, y, z};
    for (int m = 0; m < 3; m++)
        HIPCHK(c, hipMemcpyAsync(c->F.x[m], h[m], sizeof(double) * c->n1, hipMemcpyHostToDevice, c->stream));
    launch_inject(c->stream, c->lv, c->ng, c->p.icrs, c->p.jcrs, c->p.kcrs, c->F.x, 3);
    for (int n = 1; n <= c->ng; n++) launch_metrics(c->stream, c->lv[n], c->F, c->C, n == 1);
    launch_outlet_ratios(c->stream, c->lv[1], c->F, c->p.btype, c->p.bdir);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->have_grid = true;
    return 0;
}

int desalm_set_coordinates(desalm_ctx* c, const double* x, const double* y, const double* z) {
    if (!c || !x || !y || !z) return fail(c, "null argument");
    HIPCHK(c, hipSetDevice(c->device));
    const double* h[3] = {x, y, z};
    for (int m = 0; m < 3; m++)
        HIPCHK(c, hipMemcpyAsync(c->F.x[m], h[m], sizeof(double) * c->n1, hipMemcpyHostToDevice, c->stream));
    launch_inject(c->stream, c->lv, c->ng, c->p.icrs, c->p.jcrs, c->p.kcrs, c->F.x, 3);
    launch_outlet_ratios(c->stream, c->lv[1], c->F, c->p.btype, c->p.bdir);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->have_grid = true;
    return 0;
}

int desalm_set_metrics(desalm_ctx* c, const double* csi, const double* eta, const double* zet, const double* aj) {
    if (!c || !csi || !eta || !zet || !aj) return fail(c, "null argument");
    HIPCHK(c, hipSetDevice(c->device));
    Level all = c->lv[1];
    all.off = 0;
    all.n = c->nall;
    if (copy_in_level(c, c->F.csi, 3, csi, all)) return -1;
    if (copy_in_level(c, c->F.eta, 3, eta, all)) return -1;
    if (copy_in_level(c, c->F.zet, 3, zet, all)) return -1;
    if (copy_in_level(c, &c->F.aj, 1, aj, all)) return -1;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int desalm_set_wall_distance(desalm_ctx* c, const double* wd, const double* dtev) {
    if (!c || !wd) return fail(c, "null argument");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpyAsync(c->F.wd, wd, sizeof(double) * c->n1, hipMemcpyHostToDevice, c->stream));
    if (dtev) HIPCHK(c, hipMemcpyAsync(c->F.dtev, dtev, sizeof(double) * c->n1, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int desalm_set_state(desalm_ctx* c, const double* q, const double* qn, const double* qnm1, const double* xnut) {
    if (!c) return fail(c, "null argument");
    HIPCHK(c, hipSetDevice(c->device));
    const Level& L = c->lv[1];
    if (copy_in_level(c, c->F.q, 5, q, L)) return -1;
    if (copy_in_level(c, c->F.qn, 5, qn, L)) return -1;
    if (copy_in_level(c, c->F.qnm1, 5, qnm1, L)) return -1;
    if (xnut) HIPCHK(c, hipMemcpyAsync(c->F.xnut, xnut, sizeof(double) * c->n1, hipMemcpyHostToDevice, c->stream));
    for (int m = 0; m < 5; m++)
        if (c->ws.q2.q[m])
            HIPCHK(c, hipMemcpyAsync(c->ws.q2.q[m], c->F.qnm1[m], sizeof(double) * c->n1, hipMemcpyDeviceToDevice,
                                     c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int desalm_get_state(desalm_ctx* c, double* q, double* qn, double* qnm1, double* xnut) {
    if (!c) return fail(c, "null argument");
    HIPCHK(c, hipSetDevice(c->device));
    const Level& L = c->lv[1];
    if (copy_out_level(c, c->F.q, 5, q, L)) return -1;
    if (copy_out_level(c, c->F.qn, 5, qn, L)) return -1;
    if (copy_out_level(c, c->F.qnm1, 5, qnm1, L)) return -1;
    if (xnut) {
        HIPCHK(c, hipMemcpyAsync(xnut, c->F.xnut, sizeof(double) * c->n1, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    return 0;
}

// ---- learned warm start (host side; kernels in k_mlp.hip) ----
extern "C++" {
namespace {

void ws_free(WarmStart& w) {
    for (void* p : w.dev) (void)hipFree(p);
    w = WarmStart();
}

template <typename T>
T* ws_alloc(desalm_ctx* c, long n) {
    void* p = nullptr;
    if (hipMalloc(&p, (size_t)(n > 0 ? n : 1) * sizeof(T)) != hipSuccess) {
        fail(c, "warm start: out of device memory");
        return nullptr;
    }
    c->ws.dev.push_back(p);
    c->ws.bytes += (long long)(n > 0 ? n : 1) * (long long)sizeof(T);
    if (hipMemset(p, 0, (size_t)(n > 0 ? n : 1) * sizeof(T)) != hipSuccess) {
        fail(c, "warm start: hipMemset failed");
        return nullptr;
    }
    return (T*)p;
}

// scratch matrices for m rows (grown, never shrunk)
int ws_scratch(desalm_ctx* c, int m) {
    WarmStart& w = c->ws;
    if (m <= w.bcap) return 0;
    const int nh = w.nh;
    w.A0 = ws_alloc<float>(c, (long)m * (kWsIn + 1));
    w.A1 = ws_alloc<float>(c, (long)m * (nh + 1));
    w.A2 = ws_alloc<float>(c, (long)m * (nh + 1));
    w.Y = ws_alloc<float>(c, (long)m * kWsOut);
    w.Tb = ws_alloc<float>(c, (long)m * kWsTgt);
    w.dZ3 = ws_alloc<float>(c, (long)m * kWsOut);
    w.dZ2 = ws_alloc<float>(c, (long)m * nh);
    w.dZ1 = ws_alloc<float>(c, (long)m * nh);
    w.P = ws_alloc<float>(c, (long)wgrad_parts(m) * (kWsIn + 1) * nh);
    if (!w.A0 || !w.A1 || !w.A2 || !w.Y || !w.Tb || !w.dZ3 || !w.dZ2 || !w.dZ1 || !w.P) return -1;
    w.bcap = m;
    return 0;
}

// forward pass of m rows already in A0
void ws_forward(desalm_ctx* c, hipStream_t s, int m) {
    WarmStart& w = c->ws;
    const int nh = w.nh;
    launch_dense(s, w.A0, kWsIn + 1, kWsIn + 1, w.w, nh, w.A1, nh + 1, m, 1, 1);
    launch_dense(s, w.A1, nh + 1, nh + 1, w.w + w.o2(), nh, w.A2, nh + 1, m, 1, 1);
    launch_dense(s, w.A2, nh + 1, nh + 1, w.w + w.o3(), kWsOut, w.Y, kWsOut, m, 0, 0);
}

// loss partials into lp, gradient into w.g, of m rows (A0 and Tb filled)
void ws_backward(desalm_ctx* c, hipStream_t s, int m, double* lp) {
    WarmStart& w = c->ws;
    const int nh = w.nh;
    launch_ws_loss(s, w.Y, w.Tb, m, w.dZ3, lp, kWsLossParts);
    launch_wgrad(s, w.A2, nh + 1, nh + 1, w.dZ3, kWsOut, kWsOut, m, w.P, w.g + w.o3());
    launch_dense_bwd(s, w.dZ3, kWsOut, kWsOut, w.w + w.o3(), nh, w.A2, nh + 1, w.dZ2, nh, m);
    launch_wgrad(s, w.A1, nh + 1, nh + 1, w.dZ2, nh, nh, m, w.P, w.g + w.o2());
    launch_dense_bwd(s, w.dZ2, nh, nh, w.w + w.o2(), nh, w.A1, nh + 1, w.dZ1, nh, m);
    launch_wgrad(s, w.A0, kWsIn + 1, kWsIn + 1, w.dZ1, nh, nh, m, w.P, w.g);
}

// the warm start of one zone, after the time-level shift
int ws_start(desalm_ctx* c, hipStream_t s) {
    WarmStart& w = c->ws;
    const Level& L = c->lv[1];
    const long nint = (long)(L.ni - 2) * (L.nj - 2) * (L.nk - 2);
    if (w.mode == DESALM_WS_EXTRAP) {
        launch_ws_apply(s, L, c->F, w.q2, nullptr, 0, nint);
        return 0;
    }
    if (!w.have_stats) return fail(c, "learned warm start: the network has not been trained (desalm_ws_train)");
    const int chunk = 1 << 17;
    if (ws_scratch(c, chunk)) return -1;
    for (long r0 = 0; r0 < nint; r0 += chunk) {
        const int m = (int)std::min((long)chunk, nint - r0);
        launch_ws_features(s, L, c->F, w.q2, w.A0, r0, m, w.stats);
        ws_forward(c, s, m);
        launch_ws_apply(s, L, c->F, w.q2, w.Y, r0, m);
    }
    return 0;
}

}  // namespace
}  // extern "C++"

// One physical step of FAS V-cycles (mg_driver.F90:2-323) over one or more
// zones that share a stream: zones[0]'s stream carries every launch.  The
// zones meet only where the reference's ranks do: the overset exchanges
// (exchange(1) = mg_bintp_tur after des_eddy, exchange(0) = mg_bintp_mom after
// the fine pass when apply_mom) and the norm reduction of mg_residual (sums
// over all zones, divided by zone 1's interior count, mg_residual.F90:36,71-82).
// Between two exchanges the zones' work is independent, so it runs zone by
// zone.  The host synchronises once per pseudo-iteration for the norms.
extern "C++" {
static bool needs_coordinates(const desalm_ctx* c) {
    for (int b = 0; b < 4; b++)
        if (c->p.btype[b] == DESALM_BC_INFLOW || c->p.btype[b] == DESALM_BC_OUTLET_MOC) return true;
    return false;
}

// solver_daf of level n for every zone of a group, one launch per kernel
// (the zones' own outlet rows, actuator-line forcing and hub zeroing per
// zone); the launch order within each zone is solver_daf's
static void solver_daf_zones(std::vector<desalm_ctx*>& zs, const ZoneList& Z, int n, int recalc, int calc_pk,
                             bool ucn_kept) {
    desalm_ctx* c0 = zs[0];
    hipStream_t s = c0->stream;
    const int kc = c0->kc_visc[n];
    timed(c0, KID_DTAU, n, [&] { launch_dtau_z(s, Z, n, ucn_kept); });
    timed(c0, KID_RESIDUAL, n, [&] { launch_residual_z(s, Z, n, 0, calc_pk, kc); });
    for (desalm_ctx* c : zs) {
        if (n == 1 && (c->C.btype[0] == DESALM_BC_OUTLET_MOC || c->C.btype[1] == DESALM_BC_OUTLET_MOC))
            timed(c0, KID_RESIDUAL, n, [&] { launch_outlet_rhs(s, c->lv[n], c->F, c->C); });
        if (!c->alm.t.empty()) timed(c0, KID_ALM, n, [&] { launch_alm_rhs(s, c->lv[n], c->F, c->alm, c->C.ren); });
    }
    timed(c0, KID_DAF_I, n, [&] { launch_daf_z(s, Z, n, 0); });
    timed(c0, KID_DAF_J, n, [&] { launch_daf_z(s, Z, n, 1); });
    timed(c0, KID_DAF_K, n, [&] { launch_daf_z(s, Z, n, 2); });
    if (n == 1)
        timed(c0, KID_BCOND, n, [&] {
            launch_bcond_fm_z(s, Z);
            for (desalm_ctx* c : zs)
                if (!c->alm.t.empty()) launch_alm_hub(s, c->lv[1], c->F, c->alm);
        });
    if (recalc == 1) timed(c0, KID_RESIDUAL_RECALC, n, [&] { launch_residual_z(s, Z, n, 1, 0, kc); });
}

template <typename Exchange>
static int vcycles(std::vector<desalm_ctx*>& zs, bool have_table, int apply_mom, Exchange exchange,
                   double* init_resd, int* itc_out, double* resid_out, const ZoneList* zl = nullptr) {
    desalm_ctx* c0 = zs[0];
    const desalm_params& p = c0->p;
    hipStream_t s = c0->stream;
    const int ng = c0->ng;
    for (desalm_ctx* c : zs)
        if (needs_coordinates(c) && !c->have_grid)
            return fail(c0, "an inflow or characteristic-outlet face (btype 4, 5) extrapolates on the grid ratio: "
                            "call desalm_set_grid or desalm_set_coordinates first");
    for (desalm_ctx* c : zs) {
        const Level& L1 = c->lv[1];
        HIPCHK(c0, hipMemsetAsync(c->F.pivots, 0, sizeof(unsigned long long), s));
        HIPCHK(c0, hipMemsetAsync(c->F.zero_pivot, 0xff, sizeof(unsigned), s));
        c->st_zero_pivot = c->st_nonfinite = c->st_diverged = 0;
        // qnm1 <- qn, qn <- q on the fine level (mg_driver.F90:48-51), and
        // q^(n-2) <- qnm1 first when a warm start keeps it
        for (int m = 0; m < 5; m++) {
            if (c->ws.q2.q[m])
                HIPCHK(c0, hipMemcpyAsync(c->ws.q2.q[m], c->F.qnm1[m], sizeof(double) * c->n1,
                                          hipMemcpyDeviceToDevice, s));
            HIPCHK(c0, hipMemcpyAsync(c->F.qnm1[m], c->F.qn[m], sizeof(double) * c->n1, hipMemcpyDeviceToDevice, s));
            HIPCHK(c0, hipMemcpyAsync(c->F.qn[m], c->F.q[m], sizeof(double) * c->n1, hipMemcpyDeviceToDevice, s));
        }
        launch_inject(s, c->lv, ng, c->p.icrs, c->p.jcrs, c->p.kcrs, c->F.qn, 5);
        launch_inject(s, c->lv, ng, c->p.icrs, c->p.jcrs, c->p.kcrs, c->F.qnm1, 5);
        if (c->ws.mode != DESALM_WS_NONE && ws_start(c, s)) {
            if (c != c0) c0->err = c->err;
            return -1;
        }
        // outlet mass-flux rescale of the starting iterate (mg_driver.F90:64-75)
        if (c->hooks & DESALM_HOOK_CONSERVATION) launch_conservation(s, L1, c->F);
        launch_copy_qold(s, L1, c->F);
    }
    HIPCHK(c0, hipGetLastError());
    const Level& L0 = c0->lv[1];
    const double tnu = (double)((long)(L0.ni - 2) * (L0.nj - 2) * (L0.nk - 2));
    // per zone: the five L1 sums and the zero-pivot word (F.resid[0..5])
    std::vector<double> parts(6 * zs.size());
    int itc = 0;
    int rc = 0;
    // decide_recalc_rh is carried from level to level and from one
    // pseudo-iteration into the next (mg_driver.F90:144-150); the body of an
    // iteration is fixed once its starting value is
    auto body = [&](int recalc_in) {
        int decide_recalc_rh = recalc_in;
        for (int n = 1; n <= ng; n++) {
            if (n > 1) {
                if (zl)
                    timed(c0, KID_NLEVEL, n, [&] { launch_nlevel_z(s, *zl, n); });
                else
                    for (desalm_ctx* c : zs)
                        timed(c, KID_NLEVEL, n, [&] {
                            launch_nlevel(s, c->lv[n], c->lv[n - 1], c->F, c->p.icrs, c->p.jcrs, c->p.kcrs);
                        });
            }
            for (int itr = 1; itr <= p.iter[n - 1]; itr++) {
                if (n != ng) decide_recalc_rh = (itr != p.iter[n - 1]) ? 0 : 1;
                if (p.des && n == 1) {
                    if (zl)
                        timed(c0, KID_DES_EDDY, 1, [&] { launch_des_eddy_z(s, *zl, c0->p.itk, c0->kc_sa); });
                    else
                        for (desalm_ctx* c : zs)
                            timed(c, KID_DES_EDDY, 1,
                                  [&] { launch_des_eddy(s, c->lv[1], c->F, c->C, c->p.itk, c->kc_sa); });
                    if (have_table) timed(c0, KID_OVERSET, 1, [&] { exchange(1); });
                }
                const bool ucn_kept = p.des && n == 1;
                if (zl)
                    solver_daf_zones(zs, *zl, n, decide_recalc_rh, itr, ucn_kept);
                else
                    for (desalm_ctx* c : zs) solver_daf(c, n, decide_recalc_rh, itr, nullptr, ucn_kept);
                // wall functions on the j walls of zone 1 (mg_driver.F90:218-224)
                if (n == 1 && (c0->hooks & DESALM_HOOK_WALL_FUNCTIONS))
                    timed(c0, KID_BCOND, 1, [&] { launch_wall_functions(s, c0->lv[1], c0->F, c0->p.ren); });
                if (n == 1 && have_table && apply_mom) timed(c0, KID_OVERSET, 1, [&] { exchange(0); });
            }
        }
        if (zl) {
            for (int n = ng - 1; n >= 1; n--) timed(c0, KID_PROLONG, n, [&] { launch_prolong_z(s, *zl, n); });
            timed(c0, KID_BCOND, 1, [&] {
                launch_bcond_fm_z(s, *zl);
                for (desalm_ctx* c : zs)
                    if (!c->alm.t.empty()) launch_alm_hub(s, c->lv[1], c->F, c->alm);
            });
            timed(c0, KID_RESID, 1, [&] { launch_resid_and_save_z(s, *zl, 1); });
        } else {
            for (desalm_ctx* c : zs) {
                for (int n = ng - 1; n >= 1; n--) {
                    timed(c, KID_PROLONG, n, [&] {
                        launch_prolong(s, c->lv[n], c->lv[n + 1], c->F, c->C, c->p.icrs, c->p.jcrs, c->p.kcrs);
                    });
                }
                timed(c, KID_BCOND, 1, [&] { bcond(c, s, true); });
                // the pressure shift, the update norms, then qold_mg <- q for
                // the next pseudo-iteration
                timed(c, KID_RESID, 1, [&] { launch_resid_and_save(s, c->lv[1], c->F, 1, true); });
            }
        }
        return decide_recalc_rh;
    };
    // the value decide_recalc_rh leaves an iteration with, from a given start
    auto recalc_out = [&](int in) {
        int v = in;
        for (int n = 1; n <= ng; n++)
            for (int itr = 1; itr <= p.iter[n - 1]; itr++)
                if (n != ng) v = (itr != p.iter[n - 1]) ? 0 : 1;
        return v;
    };
    auto schedule = [&](int in) {
        std::vector<int> v;
        int r = in;
        for (int n = 1; n <= ng; n++)
            for (int itr = 1; itr <= p.iter[n - 1]; itr++) {
                if (n != ng) r = (itr != p.iter[n - 1]) ? 0 : 1;
                v.push_back(r);
            }
        return v;
    };
    // Without the per-launch profiler the iteration's launches are captured
    // once into a HIP graph and replayed; this needs every iteration to run the
    // same launches, i.e. the same decide_recalc_rh schedule from the start
    // value 0 of the first iteration as from the carried value of the others.
    bool profiling = false;
    for (desalm_ctx* c : zs) profiling = profiling || c->prof.on;
    const int carried = recalc_out(0);
    const bool use_graph = !profiling && p.itm > 1 && schedule(0) == schedule(carried);
    // the executable graph is destroyed on every way out of this function
    struct ExecGuard {
        hipGraphExec_t e = nullptr;
        ~ExecGuard() {
            if (e) (void)hipGraphExecDestroy(e);
        }
    } guard;
    hipGraphExec_t& exec = guard.e;
    int decide_recalc_rh = 0;
    for (int it = 1; it <= p.itm; it++) {
        itc++;
        if (use_graph) {
            if (!exec) {
                hipGraph_t graph = nullptr;
                HIPCHK(c0, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                body(decide_recalc_rh);
                const hipError_t ce = hipStreamEndCapture(s, &graph);
                if (ce != hipSuccess) return fail(c0, std::string("graph capture failed: ") + hipGetErrorString(ce));
                const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
                if (ie != hipSuccess) return fail(c0, std::string("graph instantiation failed: ") + hipGetErrorString(ie));
            }
            const hipError_t le = hipGraphLaunch(exec, s);
            if (le != hipSuccess) return fail(c0, std::string("hipGraphLaunch failed: ") + hipGetErrorString(le));
            decide_recalc_rh = carried;
        } else {
            decide_recalc_rh = body(decide_recalc_rh);
        }
        HIPCHK(c0, hipGetLastError());
        for (size_t z = 0; z < zs.size(); z++)
            HIPCHK(c0, hipMemcpyAsync(&parts[6 * z], zs[z]->F.resid, 6 * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(c0, hipStreamSynchronize(s));
        for (desalm_ctx* c : zs) prof_flush(c);
        double resd[5] = {0, 0, 0, 0, 0};
        for (size_t z = 0; z < zs.size(); z++)
            for (int m = 0; m < 5; m++) resd[m] = resd[m] + parts[6 * z + m];
        double residual[5];
        for (int m = 0; m < 5; m++) residual[m] = log10(resd[m] / tnu);
        if (resid_out)
            for (int m = 0; m < 5; m++) resid_out[5 * (it - 1) + m] = residual[m];
        // failure detection: the first row of a line solve whose pivot is
        // zero or not finite (SGTSV's INFO), an update norm that is not
        // finite, and growth of the largest u, v, w norm by more than
        // kDivergeDecades over the first iteration's
        unsigned zp = 0xffffffffu;
        for (size_t z = 0; z < zs.size(); z++) {
            unsigned w = 0;
            memcpy(&w, &parts[6 * z + 5], sizeof(unsigned));
            zp = std::min(zp, w);
        }
        bool nonfinite = false;
        for (int m = 0; m < 5; m++) nonfinite = nonfinite || !std::isfinite(resd[m]);
        // mg_ck_conver (mg_driver.F90:372-398)
        if (itc == 1)
            for (int m = 0; m < 5; m++) init_resd[m] = residual[m];
        const double eomax = fmax(fmax(residual[1], residual[2]), residual[3]);
        const double ermax = fmax(fmax(residual[1] - init_resd[1], residual[2] - init_resd[2]),
                                  residual[3] - init_resd[3]);
        if (zp != 0xffffffffu) {
            c0->st_zero_pivot = (int)zp;
            rc = (int)zp;
            c0->err = "a line solve met a zero or non-finite pivot at row " + std::to_string(zp) +
                      " in pseudo-iteration " + std::to_string(itc);
            break;
        }
        if (nonfinite) {
            c0->st_nonfinite = itc;
            rc = DESALM_ERR_NONFINITE;
            c0->err = "the update norms are not finite after pseudo-iteration " + std::to_string(itc);
            break;
        }
        if (itc > 1 && ermax > kDivergeDecades && !c0->st_diverged) c0->st_diverged = itc;
        bool converged = false;
        if (itc > p.it_min && ermax < p.er_min) converged = true;
        if (itc > p.it_min && eomax < p.eo_min) converged = true;
        if (converged) break;
    }
    for (desalm_ctx* c : zs)
        HIPCHK(c0, hipMemcpyAsync(&c->last_pivots, c->F.pivots, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    if (itc_out) *itc_out = itc;
    return rc;
}

// Entry state of a physical step that the fallback restores: qnm1, q^(n-2)
// and xnut of the fine level (q and qn come back from the shifted qn and
// qnm1, which the pseudo-iterations leave alone).
static int fb_save(desalm_ctx* c) {
    const long n1 = c->n1;
    if (!c->fb_save) {
        c->fb_save = dalloc(c, 11 * n1);
        if (!c->fb_save) return fail(c, "fallback: out of device memory");
    }
    hipStream_t s = c->stream;
    for (int m = 0; m < 5; m++) {
        HIPCHK(c, hipMemcpyAsync(c->fb_save + m * n1, c->F.qnm1[m], sizeof(double) * n1, hipMemcpyDeviceToDevice, s));
        if (c->ws.q2.q[m])
            HIPCHK(c, hipMemcpyAsync(c->fb_save + (5 + m) * n1, c->ws.q2.q[m], sizeof(double) * n1,
                                     hipMemcpyDeviceToDevice, s));
    }
    HIPCHK(c, hipMemcpyAsync(c->fb_save + 10 * n1, c->F.xnut, sizeof(double) * n1, hipMemcpyDeviceToDevice, s));
    return 0;
}

static int fb_restore(desalm_ctx* c) {
    const long n1 = c->n1;
    hipStream_t s = c->stream;
    for (int m = 0; m < 5; m++) {
        HIPCHK(c, hipMemcpyAsync(c->F.q[m], c->F.qn[m], sizeof(double) * n1, hipMemcpyDeviceToDevice, s));
        HIPCHK(c, hipMemcpyAsync(c->F.qn[m], c->F.qnm1[m], sizeof(double) * n1, hipMemcpyDeviceToDevice, s));
        HIPCHK(c, hipMemcpyAsync(c->F.qnm1[m], c->fb_save + m * n1, sizeof(double) * n1, hipMemcpyDeviceToDevice, s));
        if (c->ws.q2.q[m])
            HIPCHK(c, hipMemcpyAsync(c->ws.q2.q[m], c->fb_save + (5 + m) * n1, sizeof(double) * n1,
                                     hipMemcpyDeviceToDevice, s));
    }
    HIPCHK(c, hipMemcpyAsync(c->F.xnut, c->fb_save + 10 * n1, sizeof(double) * n1, hipMemcpyDeviceToDevice, s));
    return 0;
}
}  // extern "C++"

int desalm_mg_driver(desalm_ctx* c, int* itc_out, double* resid_out) {
    if (!c) return fail(c, "null argument");
    HIPCHK(c, hipSetDevice(c->device));
    hipStream_t s = c->stream;
    HIPCHK(c, hipEventRecord(c->ev0, s));
    std::vector<desalm_ctx*> zs{c};
    const Level& L1 = c->lv[1];
    auto exchange = [&](int tur) { launch_bintp(s, L1, c->F, c->C, c->ov, tur, c->cvin, c->p.des); };
    // the FP32-DAF form and the warm starts fall back to the plain FP64
    // method for a step they do not carry: a zero or non-finite pivot, norms
    // that are not finite, or growth past kDivergeDecades
    const bool plain = !c->C.mixed && c->ws.mode == DESALM_WS_NONE;
    if (!plain && fb_save(c)) return -1;
    int rc = vcycles(zs, c->ov.n > 0, c->ov_mom, exchange, c->init_resd, itc_out, resid_out);
    if (rc == -1) return -1;
    if (!plain && (rc != 0 || c->st_diverged)) {
        if (fb_restore(c)) return -1;
        const int mixed = c->C.mixed, mode = c->ws.mode;
        c->C.mixed = 0;
        c->ws.mode = DESALM_WS_NONE;
        rc = vcycles(zs, c->ov.n > 0, c->ov_mom, exchange, c->init_resd, itc_out, resid_out);
        c->C.mixed = mixed;
        c->ws.mode = mode;
        c->fallbacks++;
        if (rc == -1) return -1;
    }
    HIPCHK(c, hipEventRecord(c->ev1, s));
    HIPCHK(c, hipStreamSynchronize(s));
    float ms = 0.0f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    return rc;
}

int desalm_solver_daf(desalm_ctx* c, int n, int decide_recalc_rh, int decide_calc_pk, double* q, const double* qn,
                      const double* qnm1, const double* csi, const double* eta, const double* zet, const double* aj,
                      const double* xnut, double* pk, double* rh) {
    if (!c || !q || !qn || !qnm1 || !csi || !eta || !zet || !aj || !xnut || !pk || !rh)
        return fail(c, "null argument");
    if (n < 1 || n > c->ng) return fail(c, "level out of range");
    HIPCHK(c, hipSetDevice(c->device));
    const Level& L = c->lv[n];
    hipStream_t s = c->stream;
    if (copy_in_level(c, c->F.q, 4, q, L)) return -1;
    if (copy_in_level(c, c->F.qn, 4, qn, L)) return -1;
    if (copy_in_level(c, c->F.qnm1, 4, qnm1, L)) return -1;
    if (copy_in_level(c, c->F.csi, 3, csi, L)) return -1;
    if (copy_in_level(c, c->F.eta, 3, eta, L)) return -1;
    if (copy_in_level(c, c->F.zet, 3, zet, L)) return -1;
    HIPCHK(c, hipMemcpyAsync(c->F.aj + L.off, aj, sizeof(double) * L.n, hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemcpyAsync(c->F.xnut + L.off, xnut, sizeof(double) * L.n, hipMemcpyHostToDevice, s));
    if (copy_in_level(c, c->F.pk, 4, pk, L)) return -1;
    double* dq = c->dq_buf;
    HIPCHK(c, hipMemsetAsync(dq, 0, sizeof(double) * 4 * L.n, s));
    HIPCHK(c, hipMemsetAsync(c->F.pivots, 0, sizeof(unsigned long long), s));
    HIPCHK(c, hipMemsetAsync(c->F.zero_pivot, 0xff, sizeof(unsigned), s));
    solver_daf(c, n, decide_recalc_rh, decide_calc_pk, decide_recalc_rh == 1 ? nullptr : dq);
    HIPCHK(c, hipGetLastError());
    if (decide_recalc_rh != 1) {
        // rh leaves solver_daf holding dq = Mc dQ*** (rhs_diag_solver.F90:624-633)
        for (int m = 0; m < 4; m++)
            HIPCHK(c, hipMemcpyAsync(c->F.rh[m] + L.off, dq + m * L.n, sizeof(double) * L.n, hipMemcpyDeviceToDevice, s));
    }
    if (copy_out_level(c, c->F.q, 4, q, L)) return -1;
    if (copy_out_level(c, c->F.pk, 4, pk, L)) return -1;
    if (copy_out_level(c, c->F.rh, 4, rh, L)) return -1;
    HIPCHK(c, hipMemcpyAsync(&c->last_pivots, c->F.pivots, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    unsigned zp = 0xffffffffu;
    HIPCHK(c, hipMemcpyAsync(&zp, c->F.zero_pivot, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    c->st_zero_pivot = zp == 0xffffffffu ? 0 : (int)zp;
    if (c->st_zero_pivot) {
        c->err = "a line solve met a zero or non-finite pivot at row " + std::to_string(zp);
        return c->st_zero_pivot;
    }
    return 0;
}

int desalm_des_eddy(desalm_ctx* c, double* q, const double* qn5, const double* qnm1_5, const double* csi,
                    const double* eta, const double* zet, const double* aj, const double* wd, const double* dtev,
                    double* xnut) {
    if (!c || !q || !qn5 || !qnm1_5 || !csi || !eta || !zet || !aj || !wd || !dtev || !xnut)
        return fail(c, "null argument");
    HIPCHK(c, hipSetDevice(c->device));
    const Level& L = c->lv[1];
    hipStream_t s = c->stream;
    if (copy_in_level(c, c->F.q, 5, q, L)) return -1;
    HIPCHK(c, hipMemcpyAsync(c->F.qn[4], qn5, sizeof(double) * L.n, hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemcpyAsync(c->F.qnm1[4], qnm1_5, sizeof(double) * L.n, hipMemcpyHostToDevice, s));
    if (copy_in_level(c, c->F.csi, 3, csi, L)) return -1;
    if (copy_in_level(c, c->F.eta, 3, eta, L)) return -1;
    if (copy_in_level(c, c->F.zet, 3, zet, L)) return -1;
    HIPCHK(c, hipMemcpyAsync(c->F.aj, aj, sizeof(double) * L.n, hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemcpyAsync(c->F.wd, wd, sizeof(double) * L.n, hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemcpyAsync(c->F.dtev, dtev, sizeof(double) * L.n, hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemcpyAsync(c->F.xnut, xnut, sizeof(double) * L.n, hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemsetAsync(c->F.pivots, 0, sizeof(unsigned long long), s));
    launch_des_eddy(s, L, c->F, c->C, c->p.itk, c->kc_sa);
    HIPCHK(c, hipGetLastError());
    if (copy_out_level(c, c->F.q, 5, q, L)) return -1;
    HIPCHK(c, hipMemcpyAsync(xnut, c->F.xnut, sizeof(double) * L.n, hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipMemcpyAsync(&c->last_pivots, c->F.pivots, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    return 0;
}

int desalm_gtsv_batch(int n, int nsys, double* dl, double* d, double* du, double* b, int* info_out) {
    if (n < 0 || nsys < 0) return fail(nullptr, "negative size");
    if (n == 0 || nsys == 0) return 0;
    if (!dl || !d || !du || !b) return fail(nullptr, "null argument");
    const size_t bytes = sizeof(double) * (size_t)n * nsys;
    double *ddl = nullptr, *dd = nullptr, *ddu = nullptr, *db = nullptr;
    int* dinfo = nullptr;
    int rc = 0;
    if (hipMalloc(&ddl, bytes) != hipSuccess || hipMalloc(&dd, bytes) != hipSuccess ||
        hipMalloc(&ddu, bytes) != hipSuccess || hipMalloc(&db, bytes) != hipSuccess ||
        hipMalloc(&dinfo, sizeof(int) * nsys) != hipSuccess) {
        rc = fail(nullptr, "device allocation failed");
    } else if (hipMemcpy(ddl, dl, bytes, hipMemcpyHostToDevice) != hipSuccess ||
               hipMemcpy(dd, d, bytes, hipMemcpyHostToDevice) != hipSuccess ||
               hipMemcpy(ddu, du, bytes, hipMemcpyHostToDevice) != hipSuccess ||
               hipMemcpy(db, b, bytes, hipMemcpyHostToDevice) != hipSuccess) {
        rc = fail(nullptr, "gtsv: host-to-device copy failed");
    } else {
        launch_gtsv_batch(0, n, nsys, ddl, dd, ddu, db, dinfo);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
            rc = fail(nullptr, "gtsv kernel failed");
        } else if (hipMemcpy(dl, ddl, bytes, hipMemcpyDeviceToHost) != hipSuccess ||
                   hipMemcpy(d, dd, bytes, hipMemcpyDeviceToHost) != hipSuccess ||
                   hipMemcpy(du, ddu, bytes, hipMemcpyDeviceToHost) != hipSuccess ||
                   hipMemcpy(b, db, bytes, hipMemcpyDeviceToHost) != hipSuccess ||
                   (info_out && hipMemcpy(info_out, dinfo, sizeof(int) * nsys, hipMemcpyDeviceToHost) != hipSuccess)) {
            rc = fail(nullptr, "gtsv: device-to-host copy failed");
        }
    }
    // teardown: a failing free leaves nothing to recover
    (void)hipFree(ddl); (void)hipFree(dd); (void)hipFree(ddu); (void)hipFree(db); (void)hipFree(dinfo);
    return rc;
}

int desalm_get_level_field(desalm_ctx* c, int which, int n, double* out) {
    if (!c || !out) return fail(c, "null argument");
    if (n < 1 || n > c->ng) return fail(c, "level out of range");
    HIPCHK(c, hipSetDevice(c->device));
    Fields& F = c->F;
    double* one[1];
    switch (which) {
        case 0: return copy_out_level(c, F.q, 5, out, c->lv[n]);
        case 1: return copy_out_level(c, F.qn, 5, out, c->lv[n]);
        case 2: return copy_out_level(c, F.qnm1, 5, out, c->lv[n]);
        case 3: return copy_out_level(c, F.qold, 5, out, c->lv[n]);
        case 4: return copy_out_level(c, F.rh, 4, out, c->lv[n]);
        case 5: return copy_out_level(c, F.pk, 4, out, c->lv[n]);
        case 6: one[0] = F.xnut; return copy_out_level(c, one, 1, out, c->lv[n]);
        case 7: return copy_out_level(c, F.csi, 3, out, c->lv[n]);
        case 8: return copy_out_level(c, F.eta, 3, out, c->lv[n]);
        case 9: return copy_out_level(c, F.zet, 3, out, c->lv[n]);
        case 10: one[0] = F.aj; return copy_out_level(c, one, 1, out, c->lv[n]);
        case 11: one[0] = F.dtev; return copy_out_level(c, one, 1, out, c->lv[n]);
        case 12: one[0] = F.wd; return copy_out_level(c, one, 1, out, c->lv[n]);
        default: return fail(c, "unknown field");
    }
}

double desalm_last_driver_ms(const desalm_ctx* c) { return c ? c->last_ms : -1.0; }

int64_t desalm_last_pivot_events(const desalm_ctx* c) { return c ? (int64_t)c->last_pivots : -1; }

int desalm_real_residual(desalm_ctx* c, double* log10_out) {
    if (!c || !log10_out) return fail(c, "null argument");
    HIPCHK(c, hipSetDevice(c->device));
    const Level& L = c->lv[1];
    hipStream_t s = c->stream;
    launch_real_residual(s, L, c->F);
    HIPCHK(c, hipGetLastError());
    double sums[5];
    HIPCHK(c, hipMemcpyAsync(sums, c->F.resid, sizeof(sums), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    const double tnu = (double)((long)(L.ni - 2) * (L.nj - 2) * (L.nk - 2));
    for (int m = 0; m < 5; m++) log10_out[m] = log10(sqrt(sums[m] / tnu));
    return 0;
}

int desalm_probe(desalm_ctx* c, int32_t n, const int32_t* ijk, double* out) {
    if (!c || n < 0 || (n > 0 && (!ijk || !out))) return fail(c, "null argument");
    if (n == 0) return 0;
    const Level& L = c->lv[1];
    std::vector<long> node(n);
    for (int p = 0; p < n; p++) {
        const int i = ijk[3 * p], j = ijk[3 * p + 1], k = ijk[3 * p + 2];
        if (i < 1 || i > L.ni || j < 1 || j > L.nj || k < 1 || k > L.nk)
            return fail(c, "desalm_probe: point " + std::to_string(p) + " outside the fine level");
        node[p] = (i - 1) + (long)L.ni * ((j - 1) + (long)L.nj * (k - 1));
    }
    HIPCHK(c, hipSetDevice(c->device));
    hipStream_t s = c->stream;
    if (n > c->probe_cap) {
        HIPCHK(c, hipStreamSynchronize(s));
        if (c->probe_dev) (void)hipFree(c->probe_dev);
        c->probe_dev = nullptr;
        c->probe_cap = 0;
        HIPCHK(c, hipMalloc(&c->probe_dev, (sizeof(long) + 6 * sizeof(double)) * (size_t)n));
        c->probe_cap = n;
    }
    long* dnode = (long*)c->probe_dev;
    double* dout = (double*)(dnode + c->probe_cap);
    HIPCHK(c, hipMemcpyAsync(dnode, node.data(), sizeof(long) * n, hipMemcpyHostToDevice, s));
    launch_probe(s, c->F, n, dnode, dout);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(out, dout, sizeof(double) * 6 * n, hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    return 0;
}

int desalm_last_status(const desalm_ctx* c, int32_t* zero_pivot_row, int32_t* nonfinite_iteration,
                       int32_t* diverged_iteration, int64_t* fallbacks) {
    if (!c) return fail(nullptr, "null argument");
    if (zero_pivot_row) *zero_pivot_row = c->st_zero_pivot;
    if (nonfinite_iteration) *nonfinite_iteration = c->st_nonfinite;
    if (diverged_iteration) *diverged_iteration = c->st_diverged;
    if (fallbacks) *fallbacks = c->fallbacks;
    return 0;
}

int desalm_profile(desalm_ctx* c, int enable) {
    if (!c) return fail(c, "null argument");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    prof_flush(c);
    c->prof.on = enable != 0;
    memset(c->prof.ms, 0, sizeof(c->prof.ms));
    memset(c->prof.count, 0, sizeof(c->prof.count));
    return 0;
}

int desalm_profile_get(const desalm_ctx* c, int kernel, int level, int64_t* count, double* total_ms) {
    if (!c || kernel < 0 || kernel >= KID_COUNT || level < 1 || level > kMaxLevels) return -1;
    if (count) *count = c->prof.count[kernel][level];
    if (total_ms) *total_ms = c->prof.ms[kernel][level];
    return 0;
}

const char* desalm_profile_name(int kernel) {
    return (kernel >= 0 && kernel < KID_COUNT) ? kKernelNames[kernel] : nullptr;
}

const char* desalm_last_error(const desalm_ctx* c) { return c ? c->err.c_str() : g_err.c_str(); }

// ---- actuator-line model ----
static void free_alm(desalm_ctx* c) {
    for (void* p : c->alm.dev) (void)hipFree(p);
    c->alm = Alm();
    c->alm_bytes = 0;
}

int desalm_set_turbines(desalm_ctx* c, int32_t n_tur, const desalm_turbine* tur, int32_t n_polar,
                        const desalm_polar* pol, double length_scale, double vel_infty) {
    if (!c) return fail(c, "null argument");
    if (n_tur < 0 || n_polar < 0 || (n_tur > 0 && (!tur || n_polar < 1 || !pol)))
        return fail(c, "desalm_set_turbines: bad arguments");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    free_alm(c);
    if (n_tur == 0) return 0;
    const Level& L1 = c->lv[1];
    const long jk = (long)L1.nj * L1.nk;
    // validate on the host: indices inside the fine level, polar indices in
    // range, distinct spreading nodes per rotor
    for (int t = 0; t < n_tur; t++) {
        const desalm_turbine& T = tur[t];
        if (T.nblades < 1 || T.nelems < 1 || T.nangles < 1 || T.n_disk < 0 || T.n_ker < 0 || T.dtheta_interp <= 0.0)
            return fail(c, "desalm_set_turbines: bad sizes of rotor " + std::to_string(t));
        const long na = (long)T.nelems * T.nangles;
        for (long e = 0; e < na; e++) {
            const int i = T.node_indexes[e], j = T.node_indexes[e + na], k = T.node_indexes[e + 2 * na];
            if (i < 1 || i + 1 > L1.ni || j < 1 || j + 1 > L1.nj || k < 1 || k + 1 > L1.nk)
                return fail(c, "desalm_set_turbines: element cell outside the fine level");
        }
        for (long e = 0; e < jk; e++)
            if (T.polar_before[e] < 0 || T.polar_before[e] >= n_polar || T.polar_after[e] < 0 ||
                T.polar_after[e] >= n_polar)
                return fail(c, "desalm_set_turbines: polar index out of range");
        std::vector<long> kn;
        for (int e = 0; e < T.n_ker; e++) {
            const int* p = T.ker_nodes + 3 * e;
            if (p[0] < 1 || p[0] > L1.ni || p[1] < 1 || p[1] > L1.nj || p[2] < 1 || p[2] > L1.nk)
                return fail(c, "desalm_set_turbines: spreading node outside the fine level");
            kn.push_back((p[0] - 1) + (long)L1.ni * ((p[1] - 1) + (long)L1.nj * (p[2] - 1)));
        }
        std::sort(kn.begin(), kn.end());
        if (std::adjacent_find(kn.begin(), kn.end()) != kn.end())
            return fail(c, "desalm_set_turbines: a spreading node is listed twice");
        for (int e = 0; e < T.n_disk; e++)
            if (T.disk_nodes[3 * e + 1] < 1 || T.disk_nodes[3 * e + 1] > L1.nj || T.disk_nodes[3 * e + 2] < 1 ||
                T.disk_nodes[3 * e + 2] > L1.nk)
                return fail(c, "desalm_set_turbines: disk node outside the fine level");
        if (T.itur1 < 1 || T.itur2 > L1.ni) return fail(c, "desalm_set_turbines: itur1..itur2 outside the fine level");
    }
    for (int p = 0; p < n_polar; p++)
        if (pol[p].n < 1 || !pol[p].angle || !pol[p].cl || !pol[p].cd)
            return fail(c, "desalm_set_turbines: empty polar " + std::to_string(p));
    bool ok = true;
    auto up = [&](const void* src, size_t bytes) -> void* {
        void* d = nullptr;
        if (hipMalloc(&d, bytes ? bytes : 8) != hipSuccess) { ok = false; return nullptr; }
        c->alm.dev.push_back(d);
        c->alm_bytes += (long long)(bytes ? bytes : 8);
        if (bytes && src && hipMemcpy(d, src, bytes, hipMemcpyHostToDevice) != hipSuccess) ok = false;
        return d;
    };
    Alm& A = c->alm;
    A.length_scale = length_scale;
    A.vel_infty = vel_infty;
    A.jmg = L1.nj;
    A.kmg = L1.nk;
    for (int t = 0; t < n_tur && ok; t++) {
        const desalm_turbine& T = tur[t];
        const long na = (long)T.nelems * T.nangles;
        AlmTurbine D;
        D.nblades = T.nblades; D.nelems = T.nelems; D.nangles = T.nangles;
        D.itur1 = T.itur1; D.itur2 = T.itur2; D.n_disk = T.n_disk; D.n_ker = T.n_ker;
        D.tur_rad = T.tur_rad; D.hub_rad = T.hub_rad; D.tsr = T.tsr; D.angle = T.angle; D.dtheta = T.dtheta_interp;
        D.node_idx = (const int*)up(T.node_indexes, sizeof(int) * 3 * na);
        D.node_pos = (const double*)up(T.node_indexes_pos, sizeof(double) * 27 * na);
        D.disk_dat = (const double*)up(T.disk_dat, sizeof(double) * 8 * jk);
        D.pol_b = (const int*)up(T.polar_before, sizeof(int) * jk);
        D.pol_a = (const int*)up(T.polar_after, sizeof(int) * jk);
        D.disk_nodes = (const int*)up(T.disk_nodes, sizeof(int) * 3 * T.n_disk);
        D.ker_nodes = (const int*)up(T.ker_nodes, sizeof(int) * 3 * T.n_ker);
        D.ker_pos = (const double*)up(T.ker_pos, sizeof(double) * 3 * T.n_ker);
        D.ker_dat = (const double*)up(T.ker_dat, sizeof(double) * 3 * T.n_ker);
        D.forces = (double*)up(nullptr, sizeof(double) * 6 * T.nblades * T.nelems);
        A.t.push_back(D);
    }
    std::vector<int> off(n_polar), cnt(n_polar);
    std::vector<double> pa, pl, pd;
    for (int p = 0; p < n_polar; p++) {
        off[p] = (int)pa.size();
        cnt[p] = pol[p].n;
        pa.insert(pa.end(), pol[p].angle, pol[p].angle + pol[p].n);
        pl.insert(pl.end(), pol[p].cl, pol[p].cl + pol[p].n);
        pd.insert(pd.end(), pol[p].cd, pol[p].cd + pol[p].n);
    }
    A.pol.off = (const int*)up(off.data(), sizeof(int) * n_polar);
    A.pol.cnt = (const int*)up(cnt.data(), sizeof(int) * n_polar);
    A.pol.angle = (const double*)up(pa.data(), sizeof(double) * pa.size());
    A.pol.cl = (const double*)up(pl.data(), sizeof(double) * pl.size());
    A.pol.cd = (const double*)up(pd.data(), sizeof(double) * pd.size());
    if (!ok) {
        free_alm(c);
        return fail(c, "desalm_set_turbines: device allocation or copy failed");
    }
    return 0;
}

// ---- several zones coupled by overset interpolation ----
struct desalm_group {
    std::vector<desalm_ctx*> z;  // zones, borrowed; zone 1 first
    // donor table split as the reference's alltoallv splits it: gather runs
    // per source zone (send side), scatter per destination zone (receive side)
    std::vector<Overset> gather, scatter;  // one entry per zone (n = 0 if none)
    std::vector<void*> dev;
    double* buf = nullptr;
    long nd = 0;
    int mom = 0;
    double init_resd[5] = {0, 0, 0, 0, 0};
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0;
    // one launch per kernel over all zones (desalm_group_set_batched)
    int batched = 1;
    ZoneList zl;
};

// the block-descriptor table of the group's zones on the device; false when
// the zones cannot share launches (more than kMaxBatchZones zones, or FP64
// and mixed-precision zones together)
static bool group_zone_table(desalm_group* g) {
    if (!g->batched || g->z.size() < 2 || g->z.size() > (size_t)kMaxBatchZones) return false;
    for (desalm_ctx* c : g->z)
        if (c->C.mixed != g->z[0]->C.mixed) return false;
    g->zl.host.clear();
    for (desalm_ctx* c : g->z) {
        ZoneDesc d{};
        for (int n = 0; n <= kMaxLevels; n++) d.lv[n] = c->lv[n];
        d.F = c->F;
        d.C = c->C;
        const Level& L1 = c->lv[1];
        d.pfix = L1.off + (c->p.ifix - 1) + (long)L1.ni * ((c->p.jfix - 1) + (long)L1.nj * (c->p.kfix - 1));
        d.icrs = c->p.icrs;
        d.jcrs = c->p.jcrs;
        d.kcrs = c->p.kcrs;
        g->zl.host.push_back(d);
    }
    if (!g->zl.dev && hipMalloc((void**)&g->zl.dev, sizeof(ZoneDesc) * kMaxBatchZones) != hipSuccess) {
        g->zl.dev = nullptr;
        return false;
    }
    return hipMemcpy(g->zl.dev, g->zl.host.data(), sizeof(ZoneDesc) * g->zl.host.size(), hipMemcpyHostToDevice) ==
           hipSuccess;
}

static void group_free_table(desalm_group* g) {
    for (void* p : g->dev) (void)hipFree(p);
    g->dev.clear();
    g->gather.assign(g->z.size(), Overset{0, nullptr, nullptr, nullptr, nullptr, nullptr});
    g->scatter.assign(g->z.size(), Overset{0, nullptr, nullptr, nullptr, nullptr, nullptr});
    g->buf = nullptr;
    g->nd = 0;
}

int desalm_group_create(desalm_ctx* const* zones, int nz, desalm_group** out) {
    if (!zones || nz < 1 || !out) return fail(nullptr, "desalm_group_create: bad arguments");
    const desalm_ctx* a = zones[0];
    for (int z = 0; z < nz; z++) {
        const desalm_ctx* b = zones[z];
        if (!b) return fail(nullptr, "desalm_group_create: null zone");
        for (int y = 0; y < z; y++)
            if (zones[y] == b) return fail(nullptr, "desalm_group_create: a zone is listed twice");
        bool same = b->device == a->device && b->ng == a->ng && b->p.itm == a->p.itm &&
                    b->p.it_min == a->p.it_min && b->p.er_min == a->p.er_min && b->p.eo_min == a->p.eo_min &&
                    b->p.des == a->p.des && b->p.itk == a->p.itk;
        for (int n = 0; n < a->ng; n++) same = same && b->p.iter[n] == a->p.iter[n];
        if (!same)
            return fail(nullptr, "desalm_group_create: zones must share the device, the multigrid levels, iter, "
                                 "itm, it_min, er_min, eo_min, des and itk (global in the reference)");
    }
    desalm_group* g = new desalm_group();
    g->z.assign(zones, zones + nz);
    group_free_table(g);
    if (hipSetDevice(a->device) != hipSuccess || hipEventCreate(&g->ev0) != hipSuccess ||
        hipEventCreate(&g->ev1) != hipSuccess) {
        delete g;
        return fail(nullptr, "desalm_group_create: HIP event creation failed");
    }
    *out = g;
    return 0;
}

int desalm_group_set_overset(desalm_group* g, int64_t n, const int32_t* hosts, const double* coef,
                             const int32_t* nodes, int apply_mom) {
    if (!g) return fail(nullptr, "null argument");
    desalm_ctx* c0 = g->z[0];
    if (n < 0 || (n > 0 && (!hosts || !coef || !nodes))) return fail(c0, "desalm_group_set_overset: bad arguments");
    HIPCHK(c0, hipSetDevice(c0->device));
    for (desalm_ctx* c : g->z) HIPCHK(c0, hipStreamSynchronize(c->stream));
    group_free_table(g);
    g->mom = apply_mom != 0;
    if (n == 0) return 0;
    const int nz = (int)g->z.size();
    std::vector<std::vector<long>> ghost(nz), gcoef_idx(nz), snode(nz), sslot(nz);
    std::vector<long> slot_of(n);
    for (int64_t d = 0; d < n; d++) {
        const int32_t* h = hosts + 4 * d;
        const int32_t* t = nodes + 4 * d;
        if (h[3] < 1 || h[3] > nz || t[3] < 1 || t[3] > nz)
            return fail(c0, "desalm_group_set_overset: zone of donor " + std::to_string(d) + " out of range");
        const Level& A = g->z[h[3] - 1]->lv[1];
        const Level& B = g->z[t[3] - 1]->lv[1];
        if (h[0] < 1 || h[0] + 1 > A.ni || h[1] < 1 || h[1] + 1 > A.nj || h[2] < 1 || h[2] + 1 > A.nk)
            return fail(c0, "desalm_group_set_overset: host cell " + std::to_string(d) + " outside its zone");
        if (t[0] < 1 || t[0] > B.ni || t[1] < 1 || t[1] > B.nj || t[2] < 1 || t[2] > B.nk)
            return fail(c0, "desalm_group_set_overset: fringe node " + std::to_string(d) + " outside its zone");
        // gather slots are numbered in the order of the source zones
        ghost[h[3] - 1].push_back((h[0] - 1) + (long)A.ni * ((h[1] - 1) + (long)A.nj * (h[2] - 1)));
        gcoef_idx[h[3] - 1].push_back(d);
        snode[t[3] - 1].push_back((t[0] - 1) + (long)B.ni * ((t[1] - 1) + (long)B.nj * (t[2] - 1)));
    }
    long slot = 0;
    for (int z = 0; z < nz; z++)
        for (long d : gcoef_idx[z]) slot_of[d] = slot++;
    for (int64_t d = 0; d < n; d++) sslot[nodes[4 * d + 3] - 1].push_back(slot_of[d]);
    for (int z = 0; z < nz; z++) {
        std::vector<long> sorted(snode[z]);
        std::sort(sorted.begin(), sorted.end());
        if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
            return fail(c0, "desalm_group_set_overset: a fringe node of zone " + std::to_string(z + 1) +
                                " is listed twice");
    }
    auto up = [&](const void* src, size_t bytes) -> void* {
        void* p = nullptr;
        if (hipMalloc(&p, bytes ? bytes : 8) != hipSuccess) return nullptr;
        g->dev.push_back(p);
        if (bytes && hipMemcpy(p, src, bytes, hipMemcpyHostToDevice) != hipSuccess) return nullptr;
        return p;
    };
    double* buf = nullptr;
    if (hipMalloc((void**)&buf, sizeof(double) * 4 * n) != hipSuccess) {
        group_free_table(g);
        return fail(c0, "desalm_group_set_overset: out of device memory");
    }
    g->dev.push_back(buf);
    g->buf = buf;
    long goff = 0;
    for (int z = 0; z < nz; z++) {
        const long nzd = (long)ghost[z].size();
        if (nzd > 0) {
            std::vector<double> cf(3 * nzd);
            for (long e = 0; e < nzd; e++)
                for (int r = 0; r < 3; r++) cf[3 * e + r] = coef[3 * gcoef_idx[z][e] + r];
            const long* dh = (const long*)up(ghost[z].data(), sizeof(long) * nzd);
            const double* dc = (const double*)up(cf.data(), sizeof(double) * 3 * nzd);
            if (!dh || !dc) { group_free_table(g); return fail(c0, "desalm_group_set_overset: upload failed"); }
            // the gather of zone z fills slots goff .. goff + nzd - 1 (4 doubles per slot)
            g->gather[z] = Overset{nzd, dh, dc, nullptr, buf + 4 * goff, nullptr};
            goff += nzd;
        }
        const long nsd = (long)snode[z].size();
        if (nsd > 0) {
            const long* dn = (const long*)up(snode[z].data(), sizeof(long) * nsd);
            const long* ds = (const long*)up(sslot[z].data(), sizeof(long) * nsd);
            if (!dn || !ds) { group_free_table(g); return fail(c0, "desalm_group_set_overset: upload failed"); }
            g->scatter[z] = Overset{nsd, nullptr, nullptr, dn, buf, ds};
        }
    }
    g->nd = n;
    return 0;
}

int desalm_group_mg_driver(desalm_group* g, int* itc_out, double* resid_out) {
    if (!g) return fail(nullptr, "null argument");
    desalm_ctx* c0 = g->z[0];
    for (desalm_ctx* c : g->z)
        if (c->hooks)
            return fail(c0, "desalm_set_hooks is for a single zone: the reference sums the outlet flux over "
                            "every rank of every zone but zone 2 (conservation.F90:104-131)");
    HIPCHK(c0, hipSetDevice(c0->device));
    hipStream_t s = c0->stream;
    // every zone's work goes on zone 1's stream for the duration of the call
    std::vector<hipStream_t> own;
    for (desalm_ctx* c : g->z) {
        HIPCHK(c0, hipStreamSynchronize(c->stream));
        own.push_back(c->stream);
        c->stream = s;
    }
    auto exchange = [&](int tur) {
        // send side: every source zone's donor values into the shared buffer;
        // the tur pass packs one value per slot, the mom pass four
        for (size_t z = 0; z < g->z.size(); z++) {
            Overset o = g->gather[z];
            if (o.n == 0) continue;
            if (tur) o.buf = g->buf + (o.buf - g->buf) / 4;
            launch_bintp_gather(s, g->z[z]->lv[1], g->z[z]->F, o, tur);
        }
        // receive side: every destination zone's fringe nodes
        for (size_t z = 0; z < g->z.size(); z++) {
            const Overset& o = g->scatter[z];
            if (o.n == 0) continue;
            desalm_ctx* c = g->z[z];
            launch_bintp_scatter(s, c->lv[1], c->F, c->C, o, tur, c->cvin, c->p.des);
        }
    };
    const bool batch = group_zone_table(g);
    int rc = hipEventRecord(g->ev0, s) == hipSuccess ? 0 : fail(c0, "hipEventRecord failed");
    if (!rc) rc = vcycles(g->z, g->nd > 0, g->mom, exchange, g->init_resd, itc_out, resid_out, batch ? &g->zl : nullptr);
    if (!rc && hipEventRecord(g->ev1, s) != hipSuccess) rc = fail(c0, "hipEventRecord failed");
    if (hipStreamSynchronize(s) != hipSuccess && !rc) rc = fail(c0, "hipStreamSynchronize failed");
    for (size_t z = 0; z < g->z.size(); z++) g->z[z]->stream = own[z];
    if (rc) return rc;
    float ms = 0.0f;
    HIPCHK(c0, hipEventElapsedTime(&ms, g->ev0, g->ev1));
    g->last_ms = ms;
    for (desalm_ctx* c : g->z) c->last_ms = ms;
    return 0;
}

int desalm_group_destroy(desalm_group* g) {
    if (!g) return 0;
    (void)hipSetDevice(g->z[0]->device);
    for (desalm_ctx* c : g->z) (void)hipStreamSynchronize(c->stream);
    group_free_table(g);
    if (g->zl.dev) (void)hipFree(g->zl.dev);
    if (g->ev0) (void)hipEventDestroy(g->ev0);
    if (g->ev1) (void)hipEventDestroy(g->ev1);
    delete g;
    return 0;
}

int desalm_group_set_batched(desalm_group* g, int on) {
    if (!g) return fail(nullptr, "null argument");
    g->batched = on ? 1 : 0;
    return 0;
}

static void free_overset(desalm_ctx* c) {
    (void)hipFree((void*)c->ov.host);
    (void)hipFree((void*)c->ov.coef);
    (void)hipFree((void*)c->ov.node);
    (void)hipFree(c->ov.buf);
    c->ov = Overset{0, nullptr, nullptr, nullptr, nullptr, nullptr};
}

int desalm_set_overset(desalm_ctx* c, int64_t n, const int32_t* hosts, const double* coef, const int32_t* nodes,
                       int apply_mom) {
    if (!c) return fail(c, "null argument");
    if (n < 0 || (n > 0 && (!hosts || !coef || !nodes))) return fail(c, "desalm_set_overset: bad arguments");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    free_overset(c);
    c->ov_mom = apply_mom != 0;
    if (n == 0) return 0;
    const Level& L = c->lv[1];
    std::vector<long> host(n), node(n);
    for (int64_t d = 0; d < n; d++) {
        const int32_t* h = hosts + 3 * d;
        const int32_t* t = nodes + 3 * d;
        if (h[0] < 1 || h[0] + 1 > L.ni || h[1] < 1 || h[1] + 1 > L.nj || h[2] < 1 || h[2] + 1 > L.nk)
            return fail(c, "desalm_set_overset: host cell " + std::to_string(d) + " outside the fine level");
        if (t[0] < 1 || t[0] > L.ni || t[1] < 1 || t[1] > L.nj || t[2] < 1 || t[2] > L.nk)
            return fail(c, "desalm_set_overset: fringe node " + std::to_string(d) + " outside the fine level");
        host[d] = (h[0] - 1) + (long)L.ni * ((h[1] - 1) + (long)L.nj * (h[2] - 1));
        node[d] = (t[0] - 1) + (long)L.ni * ((t[1] - 1) + (long)L.nj * (t[2] - 1));
    }
    {
        std::vector<long> sorted(node);
        std::sort(sorted.begin(), sorted.end());
        if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
            return fail(c, "desalm_set_overset: a fringe node is listed twice");
    }
    long *dh = nullptr, *dn = nullptr;
    double *dc = nullptr, *db = nullptr;
    if (hipMalloc((void**)&dh, sizeof(long) * n) != hipSuccess || hipMalloc((void**)&dn, sizeof(long) * n) != hipSuccess ||
        hipMalloc((void**)&dc, sizeof(double) * 3 * n) != hipSuccess ||
        hipMalloc((void**)&db, sizeof(double) * 4 * n) != hipSuccess) {
        (void)hipFree(dh); (void)hipFree(dn); (void)hipFree(dc); (void)hipFree(db);
        return fail(c, "desalm_set_overset: out of device memory");
    }
    c->ov = Overset{(long)n, dh, dc, dn, db, nullptr};
    HIPCHK(c, hipMemcpy(dh, host.data(), sizeof(long) * n, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(dn, node.data(), sizeof(long) * n, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(dc, coef, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
    return 0;
}

int desalm_bintp(desalm_ctx* c, int which) {
    if (!c) return fail(c, "null argument");
    HIPCHK(c, hipSetDevice(c->device));
    launch_bintp(c->stream, c->lv[1], c->F, c->C, c->ov, which != 0, c->cvin, c->p.des);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int64_t desalm_device_bytes(const desalm_ctx* c) {
    if (!c) return -1;
    return c->dev_bytes + c->alm_bytes + c->ws.bytes;
}

int desalm_set_tiles(desalm_ctx* c, const int* kc_visc, int kc_sa) {
    if (!c) return fail(c, "null argument");
    for (int n = 1; n <= c->ng; n++) {
        const int v = kc_visc ? kc_visc[n - 1] : 0;
        if (v < 0 || v > 64) return fail(c, "tiles: planes per workgroup must be in 0..64");
    }
    if (kc_sa < 0 || kc_sa > 64) return fail(c, "tiles: planes per workgroup must be in 0..64");
    for (int n = 1; n <= c->ng; n++) c->kc_visc[n] = kc_visc ? kc_visc[n - 1] : 0;
    c->kc_sa = kc_sa;
    return 0;
}

int desalm_get_tiles(const desalm_ctx* c, int* kc_visc, int* kc_sa) {
    if (!c) return fail(nullptr, "null argument");
    for (int n = 1; n <= c->ng; n++)
        if (kc_visc) kc_visc[n - 1] = c->kc_visc[n] > 0 ? c->kc_visc[n] : visc_kc_default(c->lv[n]);
    if (kc_sa) *kc_sa = c->kc_sa > 0 ? c->kc_sa : sa_kc_default(c->lv[1]);
    return 0;
}

int desalm_autotune(desalm_ctx* c, int reps) {
    if (!c) return fail(c, "null argument");
    if (reps < 1 || reps > 1000) return fail(c, "autotune: reps must be in 1..1000");
    HIPCHK(c, hipSetDevice(c->device));
    hipStream_t s = c->stream;
    static const int cand[] = {1, 2, 3, 4, 6, 8, 12, 16};
    // times one launch configuration: a warm-up launch, then reps launches
    // between two events; the launches only rewrite scratch (rh, rh5, av, dtj)
    auto time_it = [&](auto launch, float& ms) -> int {
        launch();
        HIPCHK(c, hipEventRecord(c->ev0, s));
        for (int r = 0; r < reps; r++) launch();
        HIPCHK(c, hipEventRecord(c->ev1, s));
        HIPCHK(c, hipEventSynchronize(c->ev1));
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
        return 0;
    };
    for (int n = 1; n <= c->ng; n++) {
        const Level& L = c->lv[n];
        float best = 1e30f;
        int pick = 0;
        for (int kc : cand) {
            if (kc > L.nk - 2) break;
            float ms = 0.0f;
            if (time_it([&] { launch_visc_pk(s, L, c->F, c->C, n, 0, kc); }, ms)) return -1;
            if (ms < best) { best = ms; pick = kc; }
        }
        c->kc_visc[n] = pick;
    }
    if (c->p.des) {
        const Level& L = c->lv[1];
        launch_sa_prep(s, L, c->F);
        float best = 1e30f;
        int pick = 0;
        for (int kc : cand) {
            if (kc > L.nk - 2) break;
            float ms = 0.0f;
            if (time_it([&] { launch_sa_rhs(s, L, c->F, c->C, kc); }, ms)) return -1;
            if (ms < best) { best = ms; pick = kc; }
        }
        c->kc_sa = pick;
    }
    HIPCHK(c, hipStreamSynchronize(s));
    return 0;
}

int desalm_set_hooks(desalm_ctx* c, int hooks) {
    if (!c) return fail(c, "null argument");
    if (hooks & ~(DESALM_HOOK_CONSERVATION | DESALM_HOOK_WALL_FUNCTIONS)) return fail(c, "unknown hook flags");
    if (hooks && !c->have_grid)
        return fail(c, "the hooks rebuild velocities from grid derivatives: call desalm_set_grid or "
                       "desalm_set_coordinates first");
    c->hooks = hooks;
    return 0;
}

int desalm_set_convergence(desalm_ctx* c, int itm, int it_min, double er_min, double eo_min) {
    if (!c) return fail(c, "null argument");
    if (itm < 0 || it_min < 0) return fail(c, "itm and it_min must be non-negative");
    c->p.itm = itm;
    c->p.it_min = it_min;
    c->p.er_min = er_min;
    c->p.eo_min = eo_min;
    return 0;
}

static void free_rfg(desalm_ctx* c) {
    for (void* p : c->rfg_dev) (void)hipFree(p);
    c->rfg_dev.clear();
    c->rfg = Rfg{0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
}

int desalm_set_rfg(desalm_ctx* c, int32_t n_modes, const double* k, const double* omega, const double* p,
                   const double* q, const double* um, const double* m, const double* cc, const double* ts,
                   const double* ls, double time0) {
    if (!c) return fail(c, "null argument");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    free_rfg(c);
    if (n_modes == 0) return 0;
    if (n_modes < 0 || !k || !omega || !p || !q || !um || !m || !cc || !ts || !ls)
        return fail(c, "desalm_set_rfg: bad arguments");
    const long nplane = (long)c->p.jmg * c->p.kmg;
    for (long e = 0; e < nplane; e++) {
        const long j = e % c->p.jmg, kk = e / c->p.jmg;
        if (j == 0 || j == c->p.jmg - 1 || kk == 0 || kk == c->p.kmg - 1) continue;
        if (!(ts[e] > 0.0) || !(ls[e] > 0.0)) return fail(c, "desalm_set_rfg: time and length scales must be > 0");
        for (int d = 0; d < 3; d++)
            if (!(cc[3 * e + d] > 0.0)) return fail(c, "desalm_set_rfg: Reynolds-stress eigenvalues must be > 0");
    }
    bool ok = true;
    auto up = [&](const double* src, long n) -> const double* {
        void* d = nullptr;
        if (hipMalloc(&d, sizeof(double) * (size_t)n) != hipSuccess) { ok = false; return nullptr; }
        c->rfg_dev.push_back(d);
        if (hipMemcpy(d, src, sizeof(double) * (size_t)n, hipMemcpyHostToDevice) != hipSuccess) ok = false;
        return (const double*)d;
    };
    c->rfg.n = n_modes;
    c->rfg.k = up(k, 3L * n_modes);
    c->rfg.omega = up(omega, n_modes);
    c->rfg.p = up(p, 3L * n_modes);
    c->rfg.q = up(q, 3L * n_modes);
    c->rfg.um = up(um, 3 * nplane);
    c->rfg.m = up(m, 9 * nplane);
    c->rfg.c = up(cc, 3 * nplane);
    c->rfg.ts = up(ts, nplane);
    c->rfg.ls = up(ls, nplane);
    if (!ok) {
        free_rfg(c);
        return fail(c, "desalm_set_rfg: out of device memory");
    }
    c->rfg_time0 = time0;
    c->steps = 0;
    return 0;
}

int desalm_step(desalm_ctx* c, int* itc_out, double* resid_out) {
    if (!c) return fail(c, "null argument");
    // inflow forcing at time = time0 + (steps so far) delti, before the
    // V-cycles (main.F90:58, rfg_inlet.F90:107-173)
    if (c->rfg.n > 0) {
        HIPCHK(c, hipSetDevice(c->device));
        launch_rfg_inflow(c->stream, c->lv[1], c->F, c->rfg, c->rfg_time0 + (double)c->steps * c->p.delti);
        HIPCHK(c, hipGetLastError());
    }
    c->steps++;
    // blade angles (main.F90:42-49); dtheta = TSR delti length_scale / tur_rad (init.F90:1447)
    const double two_pi = 2.0 * 3.141592653589793;
    for (AlmTurbine& T : c->alm.t) {
        if (T.angle > two_pi) T.angle = T.angle - two_pi;
        T.angle = T.angle + T.tsr * c->p.delti * c->alm.length_scale / T.tur_rad;
    }
    return desalm_mg_driver(c, itc_out, resid_out);
}

int desalm_ws_configure(desalm_ctx* c, int hidden, int64_t per_record, int64_t capacity, uint64_t seed) {
    if (!c) return fail(c, "null argument");
    if (hidden < 1 || hidden > kWsMaxHidden) return fail(c, "warm start: hidden must be in 1..64");
    if (per_record < 1 || capacity < per_record) return fail(c, "warm start: need 1 <= samples_per_record <= capacity");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    ws_free(c->ws);
    WarmStart& w = c->ws;
    w.nh = hidden;
    w.per_step = per_record;
    w.cap = capacity;
    w.seed = seed;
    for (int m = 0; m < 5; m++) {
        w.q2.q[m] = ws_alloc<double>(c, c->n1);
        if (!w.q2.q[m]) return -1;
        HIPCHK(c, hipMemcpyAsync(w.q2.q[m], c->F.qnm1[m], sizeof(double) * c->n1, hipMemcpyDeviceToDevice, c->stream));
    }
    w.X = ws_alloc<float>(c, capacity * (kWsIn + 1));
    w.T = ws_alloc<float>(c, capacity * kWsTgt);
    const int np = w.np();
    w.w = ws_alloc<float>(c, np);
    w.g = ws_alloc<float>(c, np);
    w.m = ws_alloc<float>(c, np);
    w.v = ws_alloc<float>(c, np);
    w.stats = ws_alloc<float>(c, kWsStats);
    if (!w.X || !w.T || !w.w || !w.g || !w.m || !w.v || !w.stats) return -1;
    // Xavier-uniform hidden layers from a splitmix64 stream, zero output layer
    std::vector<float> h(np, 0.0f);
    unsigned long long st = seed;
    auto uni = [&]() {
        unsigned long long z = (st += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    };
    const int nh = hidden;
    const double a1 = sqrt(6.0 / (kWsIn + nh)), a2 = sqrt(6.0 / (nh + nh));
    for (int k = 0; k < kWsIn; k++)
        for (int n = 0; n < nh; n++) h[(size_t)k * nh + n] = (float)((2.0 * uni() - 1.0) * a1);
    for (int k = 0; k < nh; k++)
        for (int n = 0; n < nh; n++) h[(size_t)w.o2() + (size_t)k * nh + n] = (float)((2.0 * uni() - 1.0) * a2);
    HIPCHK(c, hipMemcpy(w.w, h.data(), sizeof(float) * np, hipMemcpyHostToDevice));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int desalm_ws_mode(desalm_ctx* c, int mode) {
    if (!c) return fail(c, "null argument");
    if (mode < DESALM_WS_NONE || mode > DESALM_WS_LEARNED) return fail(c, "warm start: unknown mode");
    if (mode != DESALM_WS_NONE && !c->ws.q2.q[0]) return fail(c, "warm start: call desalm_ws_configure first");
    c->ws.mode = mode;
    return 0;
}

int desalm_ws_record(desalm_ctx* c) {
    if (!c) return fail(c, "null argument");
    WarmStart& w = c->ws;
    if (!w.q2.q[0]) return fail(c, "warm start: call desalm_ws_configure first");
    HIPCHK(c, hipSetDevice(c->device));
    const long at = (w.records * w.per_step) % w.cap;
    const long m = std::min(w.per_step, w.cap - at);
    launch_ws_sample(c->stream, c->lv[1], c->F, w.q2, w.X + at * (kWsIn + 1), w.T + at * kWsTgt, m, w.seed,
                     w.records);
    HIPCHK(c, hipGetLastError());
    w.records++;
    w.rows = std::min(w.cap, w.rows + m);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int desalm_ws_train(desalm_ctx* c, int iters, int batch, double lr, double* loss_out) {
    if (!c) return fail(c, "null argument");
    WarmStart& w = c->ws;
    if (!w.q2.q[0]) return fail(c, "warm start: call desalm_ws_configure first");
    if (w.rows < 1) return fail(c, "warm start: no samples recorded (desalm_ws_record)");
    if (iters < 0 || batch < 1) return fail(c, "warm start: iters >= 0 and batch >= 1 required");
    HIPCHK(c, hipSetDevice(c->device));
    hipStream_t s = c->stream;
    if (ws_scratch(c, batch)) return -1;
    if (!w.have_stats) {
        // per-column mean and 1/sd of the buffer's features (a column whose
        // sd is below 1e-30 of its scale is dropped), and per component the
        // root mean square of d1 as the target scale (1 when it is zero)
        const int nparts = 64;
        double* part = ws_alloc<double>(c, (long)nparts * kWsIn * 2);
        if (!part) return -1;
        std::vector<double> hp((size_t)nparts * kWsIn * 2);
        std::vector<float> st(kWsStats, 0.0f);
        for (int pass = 0; pass < 2; pass++) {
            const int nc = pass == 0 ? kWsIn : kWsTgt;
            launch_colstats(s, pass == 0 ? w.X : w.T, pass == 0 ? kWsIn + 1 : kWsTgt, nc, w.rows, part, nparts);
            HIPCHK(c, hipGetLastError());
            HIPCHK(c, hipMemcpyAsync(hp.data(), part, sizeof(double) * nparts * nc * 2, hipMemcpyDeviceToHost, s));
            HIPCHK(c, hipStreamSynchronize(s));
            for (int col = 0; col < nc; col++) {
                double s1 = 0.0, s2 = 0.0;
                for (int b = 0; b < nparts; b++) {
                    s1 += hp[((size_t)b * nc + col) * 2];
                    s2 += hp[((size_t)b * nc + col) * 2 + 1];
                }
                const double mu = s1 / (double)w.rows;
                if (pass == 0) {
                    const double sd = sqrt(std::max(0.0, s2 / (double)w.rows - mu * mu));
                    st[col] = (float)mu;
                    st[kWsIn + col] = sd > 1e-30 * std::max(1.0, fabs(mu)) ? (float)(1.0 / sd) : 0.0f;
                } else if (col >= kWsComp && col < 2 * kWsComp) {
                    const double rms = sqrt(s2 / (double)w.rows);
                    st[2 * kWsIn + col - kWsComp] = rms > 0.0 ? (float)rms : 1.0f;
                }
            }
        }
        HIPCHK(c, hipMemcpy(w.stats, st.data(), sizeof(float) * st.size(), hipMemcpyHostToDevice));
        w.have_stats = true;
    }
    if (w.lcap < (long)iters * kWsLossParts) {
        w.lpart = ws_alloc<double>(c, (long)iters * kWsLossParts);
        if (!w.lpart) return -1;
        w.lcap = (long)iters * kWsLossParts;
    }
    const float b1 = 0.9f, b2 = 0.999f;
    for (int it = 0; it < iters; it++) {
        launch_ws_batch(s, w.X, w.T, w.rows, w.A0, w.Tb, batch, w.seed, w.adam_t, w.stats);
        ws_forward(c, s, batch);
        ws_backward(c, s, batch, w.lpart + (long)it * kWsLossParts);
        w.adam_t++;
        const float c1 = 1.0f - powf(b1, (float)w.adam_t), c2 = 1.0f - powf(b2, (float)w.adam_t);
        launch_adam(s, w.w, w.g, w.m, w.v, w.np(), (float)lr, b1, b2, 1e-8f, c1, c2);
    }
    HIPCHK(c, hipGetLastError());
    if (loss_out && iters > 0) {
        std::vector<double> lp((size_t)iters * kWsLossParts);
        HIPCHK(c, hipMemcpyAsync(lp.data(), w.lpart, sizeof(double) * lp.size(), hipMemcpyDeviceToHost, s));
        HIPCHK(c, hipStreamSynchronize(s));
        for (int it = 0; it < iters; it++) {
            double t = 0.0;
            for (int b = 0; b < kWsLossParts; b++) t += lp[(size_t)it * kWsLossParts + b];
            loss_out[it] = t / ((double)batch * kWsComp);
        }
    }
    HIPCHK(c, hipStreamSynchronize(s));
    return 0;
}

int desalm_ws_history(desalm_ctx* c, double* get, const double* set) {
    if (!c) return fail(c, "null argument");
    if (!c->ws.q2.q[0]) return fail(c, "warm start: call desalm_ws_configure first");
    HIPCHK(c, hipSetDevice(c->device));
    const Level& L = c->lv[1];
    if (copy_out_level(c, c->ws.q2.q, 5, get, L)) return -1;
    if (copy_in_level(c, c->ws.q2.q, 5, set, L)) return -1;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int desalm_turbine_angle(desalm_ctx* c, int t, double* get, const double* set) {
    if (!c) return fail(c, "null argument");
    if (t < 0 || t >= (int)c->alm.t.size()) return fail(c, "no such turbine");
    if (get) *get = c->alm.t[t].angle;
    if (set) c->alm.t[t].angle = *set;
    return 0;
}

int desalm_ws_param_count(const desalm_ctx* c) { return c && c->ws.nh > 0 ? c->ws.np() : 0; }

int desalm_ws_params(desalm_ctx* c, float* get, const float* set) {
    if (!c) return fail(c, "null argument");
    if (!c->ws.w) return fail(c, "warm start: call desalm_ws_configure first");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (get) HIPCHK(c, hipMemcpy(get, c->ws.w, sizeof(float) * c->ws.np(), hipMemcpyDeviceToHost));
    if (set) HIPCHK(c, hipMemcpy(c->ws.w, set, sizeof(float) * c->ws.np(), hipMemcpyHostToDevice));
    return 0;
}

int desalm_ws_stats(desalm_ctx* c, float* get, const float* set) {
    if (!c) return fail(c, "null argument");
    if (!c->ws.stats) return fail(c, "warm start: call desalm_ws_configure first");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const size_t n = kWsStats;
    if (get) HIPCHK(c, hipMemcpy(get, c->ws.stats, sizeof(float) * n, hipMemcpyDeviceToHost));
    if (set) {
        HIPCHK(c, hipMemcpy(c->ws.stats, set, sizeof(float) * n, hipMemcpyHostToDevice));
        c->ws.have_stats = true;
    }
    return 0;
}

int desalm_ws_features(desalm_ctx* c, int64_t m, float* x, float* t) {
    if (!c || !x) return fail(c, "null argument");
    WarmStart& w = c->ws;
    if (!w.q2.q[0]) return fail(c, "warm start: call desalm_ws_configure first");
    const Level& L = c->lv[1];
    if (m < 0 || m > (long)(L.ni - 2) * (L.nj - 2) * (L.nk - 2)) return fail(c, "warm start: m out of range");
    HIPCHK(c, hipSetDevice(c->device));
    hipStream_t s = c->stream;
    float* X = nullptr;
    HIPCHK(c, hipMalloc(&X, sizeof(float) * (size_t)std::max<int64_t>(m, 1) * (kWsIn + 1)));
    // raw rows: the inference features kernel without statistics
    launch_ws_features(s, L, c->F, w.q2, X, 0, m, nullptr);
    std::vector<float> hx((size_t)m * (kWsIn + 1));
    const hipError_t e1 = hipMemcpyAsync(hx.data(), X, sizeof(float) * hx.size(), hipMemcpyDeviceToHost, s);
    const hipError_t e2 = hipStreamSynchronize(s);
    (void)hipFree(X);
    if (e1 != hipSuccess || e2 != hipSuccess) return fail(c, "warm start: feature read-back failed");
    for (long r = 0; r < m; r++)
        for (int f = 0; f < kWsIn; f++) x[r * kWsIn + f] = hx[(size_t)r * (kWsIn + 1) + f];
    if (t) {
        // targets of the same ordinals: r = q - 2 q^n + q^(n-1), d1, d2
        std::vector<double> q((size_t)c->n1), qn((size_t)c->n1), qm((size_t)c->n1), q2((size_t)c->n1);
        for (int comp = 0; comp < kWsComp; comp++) {
            HIPCHK(c, hipMemcpy(q.data(), c->F.q[comp], sizeof(double) * c->n1, hipMemcpyDeviceToHost));
            HIPCHK(c, hipMemcpy(qn.data(), c->F.qn[comp], sizeof(double) * c->n1, hipMemcpyDeviceToHost));
            HIPCHK(c, hipMemcpy(qm.data(), c->F.qnm1[comp], sizeof(double) * c->n1, hipMemcpyDeviceToHost));
            HIPCHK(c, hipMemcpy(q2.data(), w.q2.q[comp], sizeof(double) * c->n1, hipMemcpyDeviceToHost));
            const long ni = L.ni - 2, nj = L.nj - 2;
            for (long r = 0; r < m; r++) {
                const long i = 1 + r % ni, j = 1 + (r / ni) % nj, k = 1 + r / (ni * nj);
                const long g = i + (long)L.ni * (j + (long)L.nj * k);
                t[r * kWsTgt + comp] = (float)(q[g] - 2.0 * qn[g] + qm[g]);
                t[r * kWsTgt + kWsComp + comp] = (float)(qn[g] - qm[g]);
                t[r * kWsTgt + 2 * kWsComp + comp] = (float)(qm[g] - q2[g]);
            }
        }
    }
    return 0;
}

int desalm_mlp_forward(desalm_ctx* c, const float* x, int64_t m, float* y) {
    if (!c || !x || !y) return fail(c, "null argument");
    WarmStart& w = c->ws;
    if (!w.w) return fail(c, "warm start: call desalm_ws_configure first");
    if (m < 1 || m > (1 << 24)) return fail(c, "mlp: m must be in 1..2^24");
    HIPCHK(c, hipSetDevice(c->device));
    hipStream_t s = c->stream;
    if (ws_scratch(c, (int)m)) return -1;
    std::vector<float> a((size_t)m * (kWsIn + 1));
    for (long r = 0; r < m; r++) {
        for (int f = 0; f < kWsIn; f++) a[(size_t)r * (kWsIn + 1) + f] = x[r * kWsIn + f];
        a[(size_t)r * (kWsIn + 1) + kWsIn] = 1.0f;
    }
    HIPCHK(c, hipMemcpyAsync(w.A0, a.data(), sizeof(float) * a.size(), hipMemcpyHostToDevice, s));
    ws_forward(c, s, (int)m);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(y, w.Y, sizeof(float) * m * kWsOut, hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    return 0;
}

int desalm_mlp_loss_grad(desalm_ctx* c, const float* x, const float* t, int64_t m, float* grad, double* loss) {
    if (!c || !x || !t || !grad) return fail(c, "null argument");
    WarmStart& w = c->ws;
    if (!w.w) return fail(c, "warm start: call desalm_ws_configure first");
    if (m < 1 || m > (1 << 24)) return fail(c, "mlp: m must be in 1..2^24");
    HIPCHK(c, hipSetDevice(c->device));
    hipStream_t s = c->stream;
    if (ws_scratch(c, (int)m)) return -1;
    if (w.lcap < kWsLossParts) {
        w.lpart = ws_alloc<double>(c, kWsLossParts);
        if (!w.lpart) return -1;
        w.lcap = kWsLossParts;
    }
    double* lp = w.lpart;
    std::vector<float> a((size_t)m * (kWsIn + 1));
    for (long r = 0; r < m; r++) {
        for (int f = 0; f < kWsIn; f++) a[(size_t)r * (kWsIn + 1) + f] = x[r * kWsIn + f];
        a[(size_t)r * (kWsIn + 1) + kWsIn] = 1.0f;
    }
    HIPCHK(c, hipMemcpyAsync(w.A0, a.data(), sizeof(float) * a.size(), hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemcpyAsync(w.Tb, t, sizeof(float) * m * kWsTgt, hipMemcpyHostToDevice, s));
    ws_forward(c, s, (int)m);
    ws_backward(c, s, (int)m, lp);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(grad, w.g, sizeof(float) * w.np(), hipMemcpyDeviceToHost, s));
    double hl[kWsLossParts];
    HIPCHK(c, hipMemcpyAsync(hl, lp, sizeof(hl), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    if (loss) {
        double tsum = 0.0;
        for (int b = 0; b < kWsLossParts; b++) tsum += hl[b];
        *loss = tsum / ((double)m * kWsComp);
    }
    return 0;
}

int desalm_finalize(desalm_ctx* c) {
    if (!c) return 0;
    // teardown: errors here leave nothing to recover, so they are ignored
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_overset(c);
    free_alm(c);
    free_rfg(c);
    ws_free(c->ws);
    prof_flush(c);
    for (hipEvent_t e : c->prof.pool) (void)hipEventDestroy(e);
    for (void* p : c->allocs) (void)hipFree(p);
    if (c->probe_dev) (void)hipFree(c->probe_dev);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

}  // extern "C"
