// k_mg.hip -- FAS multigrid transfers, fine-grid boundary conditions,
// update norms and layout conversion.
//
//   k_inject        mg_inject (src/mg_inject.F90:2-64)
//   k_nlevel        mg_nlevel: injection of q, xnut, qold_mg and the 27-point
//                   restriction of rh into pk (src/mg_nlevel.F90:2-150),
//                   boundary pk zeroed (mg_zero_pk_boundary)
//   k_prolong_fill  mg_prolong's coarse correction q - qold_mg and its
//                   k/j/i linear fills onto the fine level
//                   (src/mg_prolong.F90:57-170), one coarse cell per thread
//   k_rs_sweep<D>   rs_sweep constant-coefficient smoothing sweeps
//                   (src/rs_sweep.F90:59-178) with every row on chip,
//                   k sweep also adds dq to q
//   k_bc_fm<D>      bcond_fm faces (src/bcond_fm.F90:79-306), k_pfix_*
//                   the pinned-pressure shift (:310-322)
//   k_resid_*       mg_residual's L1 update norms (src/mg_residual.F90:58-67)
#include "desalm_internal.h"
#include "seg_sweep.h"

namespace desalm {

namespace {

__device__ __forceinline__ long lidx(const Level& L, int i, int j, int k) {
    return (long)i + (long)L.ni * ((long)j + (long)L.nj * k);
}

// coarse node (i,j,k) -> coincident fine node, 0-based (fine 1-based index
// 2^crs * i - crs of src/mg_nlevel.F90:22-24)
__device__ __forceinline__ int c2f(int i, int crs) { return crs ? 2 * i : i; }

__global__ void k_inject(Level Lc, Level Lf, double* const* comps, int ncomp, int icrs, int jcrs, int kcrs) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= Lc.n) return;
    const int i = (int)(t % Lc.ni), j = (int)((t / Lc.ni) % Lc.nj), k = (int)(t / ((long)Lc.ni * Lc.nj));
    const long gf = Lf.off + lidx(Lf, c2f(i, icrs), c2f(j, jcrs), c2f(k, kcrs));
    const long gc = Lc.off + t;
    for (int m = 0; m < ncomp; m++) comps[m][gc] = comps[m][gf];
}

struct Ptrs {
    double* p[16];
};

__global__ void k_inject_p(Level Lc, Level Lf, Ptrs P, int ncomp, int icrs, int jcrs, int kcrs) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= Lc.n) return;
    const int i = (int)(t % Lc.ni), j = (int)((t / Lc.ni) % Lc.nj), k = (int)(t / ((long)Lc.ni * Lc.nj));
    const long gf = Lf.off + lidx(Lf, c2f(i, icrs), c2f(j, jcrs), c2f(k, kcrs));
    const long gc = Lc.off + t;
    for (int m = 0; m < ncomp; m++) P.p[m][gc] = P.p[m][gf];
}

__device__ __forceinline__ void nlevel_body(const Level& Lc, const Level& Lf, const Fields& F, int icrs, int jcrs,
                                            int kcrs, unsigned bid, unsigned nb) {
    const long t = (long)xcd_block(bid, nb) * blockDim.x + threadIdx.x;
    if (t >= Lc.n) return;
    const int ni = Lc.ni, nj = Lc.nj, nk = Lc.nk;
    const int i = (int)(t % ni), j = (int)((t / ni) % nj), k = (int)(t / ((long)ni * nj));
    const long ll = Lf.off + lidx(Lf, c2f(i, icrs), c2f(j, jcrs), c2f(k, kcrs));
    const long l = Lc.off + t;
    // loads before stores throughout, so that they are in flight together
    double v[4];
#pragma unroll
    for (int m = 0; m < 4; m++) v[m] = F.q[m][ll];
    const double xn = F.xnut[ll];
#pragma unroll
    for (int m = 0; m < 4; m++) {
        F.q[m][l] = v[m];
        F.qold[m][l] = v[m];
    }
    F.xnut[l] = xn;
    if (i == 0 || i == ni - 1 || j == 0 || j == nj - 1 || k == 0 || k == nk - 1) {
        // the reference evaluates these from out-of-level neighbours and then
        // zeroes them (src/mg_driver.F90:134)
#pragma unroll
        for (int m = 0; m < 4; m++) F.pk[m][l] = 0.0;
        return;
    }
    const long ic = 1, ie = Lf.ni, iz = (long)Lf.ni * Lf.nj;
    double pkv[4];
#pragma unroll
    for (int m = 0; m < 4; m++) {
        const double* r = F.rh[m] + ll;
        pkv[m] = (0.25 * r[ic + ie + iz] + 0.5 * r[ic + ie] + 0.25 * r[ic + ie - iz] +
                   0.5 * r[ie + iz] + r[ie] + 0.5 * r[ie - iz] +
                   0.25 * r[-ic + ie + iz] + 0.5 * r[-ic + ie] + 0.25 * r[-ic + ie - iz] +
                   0.5 * r[ic + iz] + r[ic] + 0.5 * r[ic - iz] +
                   r[iz] + 2.0 * r[0] + r[-iz] +
                   0.5 * r[-ic + iz] + r[-ic] + 0.5 * r[-ic - iz] +
                   0.25 * r[ic - ie + iz] + 0.5 * r[ic - ie] + 0.25 * r[ic - ie - iz] +
                   0.5 * r[-ie + iz] + r[-ie] + 0.5 * r[-ie - iz] +
                   0.25 * r[-ic - ie + iz] + 0.5 * r[-ic - ie] + 0.25 * r[-ic - ie - iz]) *
                  (1.0 / 16.0);
    }
#pragma unroll
    for (int m = 0; m < 4; m++) F.pk[m][l] = pkv[m];
}

__global__ void k_zero_pk_boundary(Level L, Fields F) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= L.n) return;
    const int ni = L.ni, nj = L.nj, nk = L.nk;
    const int i = (int)(t % ni), j = (int)((t / ni) % nj), k = (int)(t / ((long)ni * nj));
    if (i == 0 || i == ni - 1 || j == 0 || j == nj - 1 || k == 0 || k == nk - 1) {
#pragma unroll
        for (int m = 0; m < 4; m++) F.pk[m][L.off + t] = 0.0;
    }
}

// ---- prolongation ----
struct ProlongCtx {
    Level Lf, Lc;
    int icrs, jcrs, kcrs;
};

// cells of the fill per direction: a coarsened direction has one cell per
// pair of adjacent coarse nodes (a, a + 1), with the fine nodes 2a (on coarse
// node a; interior for a >= 1) and 2a + 1 (between a and a + 1); a direction
// that is not coarsened has one cell per interior fine node f = a + 1, on
// coarse node a + 1.  Coarsened directions have odd fine node counts, so
// 2a + 1 <= nf - 2 for every cell.
__host__ __device__ inline int fill_cells(int nf, int nc, int crs) { return crs ? nc - 1 : nf - 2; }
__host__ __device__ inline long fill_cells(const ProlongCtx& P) {
    return (long)fill_cells(P.Lf.ni, P.Lc.ni, P.icrs) * fill_cells(P.Lf.nj, P.Lc.nj, P.jcrs) *
           fill_cells(P.Lf.nk, P.Lc.nk, P.kcrs);
}

// The coarse correction q - qold_mg (mg_prolong.F90:57-75) filled linearly
// onto the fine level's interior nodes into F.dp, level-local: k, then j,
// then i (mg_prolong.F90:82-161).  One thread per cell: it loads q and
// qold_mg at the cell's eight coarse corners once and forms each of its (up
// to eight) fine nodes from them.  A fine node coincident with a coarse node
// in a direction takes the low corner there (the reference's high and low
// nodes are then the same node), one between two coarse nodes takes half of
// each: the same operands in the same order as the reference's fills.
__device__ __forceinline__ void prolong_fill_body(const ProlongCtx& P, const Fields& F, unsigned bid, unsigned nb) {
    const Level& Lf = P.Lf;
    const Level& Lc = P.Lc;
    const int ca = fill_cells(Lf.ni, Lc.ni, P.icrs), cb = fill_cells(Lf.nj, Lc.nj, P.jcrs);
    const long t = (long)xcd_block(bid, nb) * blockDim.x + threadIdx.x;
    if (t >= fill_cells(P)) return;
    const int a = (int)(t % ca), b = (int)((t / ca) % cb), c = (int)(t / ((long)ca * cb));
    // per direction: low and high coarse corner, fine index of node 0, and
    // whether nodes 0 (coincident) and 1 (between) exist
    const int cl[3] = {P.icrs ? a : a + 1, P.jcrs ? b : b + 1, P.kcrs ? c : c + 1};
    const int crs[3] = {P.icrs, P.jcrs, P.kcrs};
    int ch[3], f0[3];
    bool has0[3], has1[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        ch[d] = crs[d] ? cl[d] + 1 : cl[d];
        f0[d] = crs[d] ? 2 * cl[d] : cl[d];
        has0[d] = !crs[d] || cl[d] >= 1;
        has1[d] = crs[d];
    }
    long corner[2][2][2];  // [i side][j side][k side], side 1 = high
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++)
#pragma unroll
            for (int z = 0; z < 2; z++)
                corner[x][y][z] = Lc.off + lidx(Lc, x ? ch[0] : cl[0], y ? ch[1] : cl[1], z ? ch[2] : cl[2]);
    const long base = lidx(Lf, f0[0], f0[1], f0[2]);
    const long sI = 1, sJ = Lf.ni, sK = (long)Lf.ni * Lf.nj;
#pragma unroll
    for (int m = 0; m < 4; m++) {
        double v[2][2][2];
#pragma unroll
        for (int x = 0; x < 2; x++)
#pragma unroll
            for (int y = 0; y < 2; y++)
#pragma unroll
                for (int z = 0; z < 2; z++) {
                    const long l = corner[x][y][z];
                    v[x][y][z] = F.q[m][l] - F.qold[m][l];
                }
        // fk[x][y][pk]: the k fill of node parity pk on the (x, y) corner column
        double fk[2][2][2];
#pragma unroll
        for (int x = 0; x < 2; x++)
#pragma unroll
            for (int y = 0; y < 2; y++) {
                fk[x][y][0] = v[x][y][0];
                fk[x][y][1] = 0.5 * v[x][y][1] + 0.5 * v[x][y][0];
            }
#pragma unroll
        for (int pk = 0; pk < 2; pk++) {
            if (!(pk ? has1[2] : has0[2])) continue;
#pragma unroll
            for (int pj = 0; pj < 2; pj++) {
                if (!(pj ? has1[1] : has0[1])) continue;
                double fj[2];
#pragma unroll
                for (int x = 0; x < 2; x++) fj[x] = pj ? 0.5 * fk[x][1][pk] + 0.5 * fk[x][0][pk] : fk[x][0][pk];
#pragma unroll
                for (int pi = 0; pi < 2; pi++) {
                    if (!(pi ? has1[0] : has0[0])) continue;
                    F.dp[m][base + pi * sI + pj * sJ + pk * sK] = pi ? 0.5 * fj[1] + 0.5 * fj[0] : fj[0];
                }
            }
        }
    }
}

// rs_sweep in direction D on the interior of a level (src/rs_sweep.F90:59-178):
// the constant system (-ep, 1+2ep, -ep) applied to the four components of dq
// (F.dp, level-local) with every row on chip (seg_sweep.h), solved in place;
// D = 2 adds the result to q instead.  The reference calls SGTSV with four
// right-hand sides; this is its non-interchanging elimination.
template <int D, int R>
__device__ __forceinline__ void rs_sweep_body(const Level& L, const Fields& F, const SegShape& sh, double epn,
                                              unsigned bid) {
    extern __shared__ double smem_rs[];
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
    double a[R], b[R], c[R], d[R][4];
#pragma unroll
    for (int r = 0; r < R; r++) {
        a[r] = -epn;
        b[r] = 1.0 + 2.0 * epn;
        c[r] = -epn;
    }
    if (D == 0) {
        // i lines: the R rows are one run of consecutive nodes, loaded as
        // such (16-byte loads; lanes are R rows apart).  A partial last
        // segment loads the line's last R rows and shifts them into place.
        const int pb = max(0, min(s0, n - R));
        const int shift = s0 - pb;
        double w[R][4];
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
            for (int m = 0; m < 4; m++) w[r][m] = F.dp[m][base + pb + r + 1];
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
            for (int m = 0; m < 4; m++) {
                double v = w[r][m];
#pragma unroll
                for (int t = 1; t < R; t++)
                    if (shift == t) v = r + t < R ? w[r + t < R ? r + t : R - 1][m] : 0.0;
                d[r][m] = r < nr ? v : 0.0;
            }
    } else {
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
            for (int m = 0; m < 4; m++) d[r][m] = r < nr ? F.dp[m][base + (long)(s0 + r + 1) * sD] : 0.0;
    }
    unsigned piv = 0;
    int zrow = 0;
    seg_solve<R, 4, double>(a, b, c, d, -epn, nr, n, s0, smem_rs, P, TI, sh.nlines, line0, li, ps, piv, zrow);
    if (D == 2) {
        // every q of the segment is loaded, and every sum formed, before
        // the first store: the loads are in flight together and the stores
        // wait for nothing
        double qv[R][4];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const long g = L.off + base + (long)(min(s0 + r, n - 1) + 1) * sD;
#pragma unroll
            for (int m = 0; m < 4; m++) qv[r][m] = F.q[m][g];
        }
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
            for (int m = 0; m < 4; m++) qv[r][m] = qv[r][m] + d[r][m];
#pragma unroll
        for (int r = 0; r < R; r++) {
            if (r < nr) {
                const long g = L.off + base + (long)(s0 + r + 1) * sD;
#pragma unroll
                for (int m = 0; m < 4; m++) F.q[m][g] = qv[r][m];
            }
        }
    } else if (D == 0 && nr == R) {
        // a whole segment of an i line: one run of consecutive nodes per
        // component
#pragma unroll
        for (int m = 0; m < 4; m++)
#pragma unroll
            for (int r = 0; r < R; r++) F.dp[m][base + s0 + r + 1] = d[r][m];
    } else {
#pragma unroll
        for (int r = 0; r < R; r++) {
            if (r < nr) {
                const long tl = base + (long)(s0 + r + 1) * sD;
#pragma unroll
                for (int m = 0; m < 4; m++) F.dp[m][tl] = d[r][m];
            }
        }
    }
    if (piv) atomicAdd(F.pivots, (unsigned long long)piv);
    if (zrow) atomicMin(F.zero_pivot, (unsigned)zrow);
}

// ---- grid ratios of the inflow faces (init_outlet_bc, init.F90:2750-2891) ----
// Face b (0..3: i = 1, i = img, j = 1, j = jmg), node t of the face plane
// (i faces: j + nj k, j faces: i + ni k): the spacing of the first node pair
// off the face over that of the second, along coordinate dir (bdir - 1).
__global__ void k_outlet_ratio(Level L, Fields F, int b, int dir) {
    const int ni = L.ni, nj = L.nj, nk = L.nk;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long np = b < 2 ? (long)nj * nk : (long)ni * nk;
    if (t >= np) return;
    const long sJ = ni, sK = (long)ni * nj;
    long l0, st;
    if (b < 2) {
        const long base = sJ * (t % nj) + sK * (t / nj);
        l0 = b == 0 ? base : base + ni - 1;
        st = b == 0 ? 1 : -1;
    } else {
        const long base = (t % ni) + sK * (t / ni);
        l0 = b == 2 ? base : base + sJ * (nj - 1);
        st = b == 2 ? sJ : -sJ;
    }
    const double* c = F.x[dir] + L.off;
    const double c0 = c[l0], c1 = c[l0 + st], c2 = c[l0 + 2 * st];
    F.rat[b][t] = (b % 2 == 0) ? (c1 - c0) / (c2 - c1) : (c0 - c1) / (c1 - c2);
}

// ---- bcond_fm faces of direction D (both sides) ----
// Inflow (btype 4) on a k face uses ratio 0: the reference fills its table
// only where btype == 4 .and. btype == 5 (init.F90:2855, :2874), i.e. never,
// and the table is zero as allocated.
// the four components of q at node g; each face below loads all it reads
// before its first store, so that the loads are in flight together
__device__ __forceinline__ void load4(const Fields& F, long g, double v[4]) {
#pragma unroll
    for (int m = 0; m < 4; m++) v[m] = F.q[m][g];
}
__device__ __forceinline__ void store4(const Fields& F, long g, const double v[4]) {
#pragma unroll
    for (int m = 0; m < 4; m++) F.q[m][g] = v[m];
}

template <int D>
__device__ __forceinline__ void bc_fm_body(const Level& L, const Fields& F, const Consts& C, unsigned bid) {
    const int ni = L.ni, nj = L.nj, nk = L.nk;
    const long sI = 1, sJ = ni, sK = (long)ni * nj;
    const long t = (long)bid * blockDim.x + threadIdx.x;
    long base, sD;
    int nd;
    if (D == 0) {
        if (t >= (long)nj * nk) return;
        base = sJ * (t % nj) + sK * (t / nj);
        sD = sI;
        nd = ni;
    } else if (D == 1) {
        if (t >= (long)ni * nk) return;
        base = (t % ni) + sK * (t / ni);
        sD = sJ;
        nd = nj;
    } else {
        if (t >= (long)ni * nj) return;
        base = t;
        sD = sK;
        nd = nk;
    }
    const long lo = L.off + base, hi = L.off + base + (long)(nd - 1) * sD;
    const int b0 = 2 * D, b1 = 2 * D + 1;
    double a[4], b[4];
    switch (C.btype[b0]) {
        case 1: case 2: case 3:
            load4(F, lo + sD, a);
            load4(F, lo + 2 * sD, b);
#pragma unroll
            for (int m = 0; m < 4; m++) a[m] = C.sa[b0][m] * a[m] + C.sb[b0][m] * b[m];
            store4(F, lo, a);
            break;
        case 6:  // plane copy from the opposite boundary plane (not done for i)
            if (D > 0) {
                load4(F, hi, a);
                store4(F, lo, a);
            }
            break;
        case 7:  // periodic wrap: the image of the opposite interior plane
            load4(F, hi - sD, a);
            store4(F, lo, a);
            break;
        case 4: {  // inflow: pressure extrapolated on the grid ratio (bcond_fm.F90:95-102)
            const double r = D < 2 ? F.rat[b0][t] : 0.0;
            F.q[0][lo] = (1.0 + r) * F.q[0][lo + sD] - r * F.q[0][lo + 2 * sD];
            break;
        }
        default: break;
    }
    switch (C.btype[b1]) {
        case 1: case 2: case 3:
            load4(F, hi - sD, a);
            load4(F, hi - 2 * sD, b);
#pragma unroll
            for (int m = 0; m < 4; m++) a[m] = C.sa[b1][m] * a[m] + C.sb[b1][m] * b[m];
            store4(F, hi, a);
            break;
        case 6:
            if (D > 0) {
                load4(F, lo, a);
                store4(F, hi, a);
            }
            break;
        case 7:
            load4(F, lo + sD, a);
            store4(F, hi, a);
            break;
        case 4: {
            const double r = D < 2 ? F.rat[b1][t] : 0.0;
            F.q[0][hi] = (1.0 + r) * F.q[0][hi - sD] - r * F.q[0][hi - 2 * sD];
            break;
        }
        case 8:
            if (D == 0) {
                load4(F, hi - sD, a);
                store4(F, hi, a);
            }
            break;
        default: break;
    }
}


__device__ __forceinline__ void pfix_sub_body(const Level& L, const Fields& F, unsigned bid) {
    const long t = (long)bid * blockDim.x + threadIdx.x;
    if (t >= L.n) return;
    const double pf = *F.pfix;
    F.q[0][L.off + t] = F.q[0][L.off + t] - pf;
}

// ---- mg_residual: deterministic two-stage L1 sums ----
// pfix: first subtract the pinned pressure F.pfix from p at every node (the
// last step of bcond_fm, deferred to this pass when it follows directly)
__device__ __forceinline__ void resid_part_body(const Level& L, const Fields& F, int save, int pfix, unsigned bid,
                                                unsigned nb) {
    __shared__ double sh[5][256];
    double acc[5] = {0, 0, 0, 0, 0};
    const int ni = L.ni, nj = L.nj, nk = L.nk;
    const double pf = pfix ? *F.pfix : 0.0;
    for (long t = (long)bid * blockDim.x + threadIdx.x; t < L.n; t += (long)nb * blockDim.x) {
        const int i = (int)(t % ni), j = (int)((t / ni) % nj), k = (int)(t / ((long)ni * nj));
        const long g = L.off + t;
        const bool in = !(i == 0 || i == ni - 1 || j == 0 || j == nj - 1 || k == 0 || k == nk - 1);
        // all ten loads before the first store
        double qv[5], qo[5];
#pragma unroll
        for (int m = 0; m < 5; m++) {
            qv[m] = F.q[m][g];
            qo[m] = F.qold[m][g];
        }
        if (pfix) {
            qv[0] = qv[0] - pf;
            F.q[0][g] = qv[0];
        }
#pragma unroll
        for (int m = 0; m < 5; m++) {
            if (in) acc[m] += fabs(qv[m] - qo[m]);
            if (save) F.qold[m][g] = qv[m];
        }
    }
#pragma unroll
    for (int m = 0; m < 5; m++) sh[m][threadIdx.x] = acc[m];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int m = 0; m < 5; m++) sh[m][threadIdx.x] += sh[m][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int m = 0; m < 5; m++) F.resid_part[bid * 5 + m] = sh[m][0];
    }
}

__device__ __forceinline__ void resid_final_body(const Fields& F, int nblocks) {
    __shared__ double sh[5][256];
    double acc[5] = {0, 0, 0, 0, 0};
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x)
#pragma unroll
        for (int m = 0; m < 5; m++) acc[m] += F.resid_part[b * 5 + m];
#pragma unroll
    for (int m = 0; m < 5; m++) sh[m][threadIdx.x] = acc[m];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int m = 0; m < 5; m++) sh[m][threadIdx.x] += sh[m][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int m = 0; m < 5; m++) F.resid[m] = sh[m][0];
    }
}

__global__ void __launch_bounds__(256) k_resid_final(Fields F, int nblocks) { resid_final_body(F, nblocks); }

__global__ void k_copy_qold(Level L, Fields F) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= L.n) return;
#pragma unroll
    for (int m = 0; m < 5; m++) F.qold[m][L.off + t] = F.q[m][L.off + t];
}

__global__ void k_aos_to_soa(const double* aos, int ncomp, long n, Ptrs P) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    for (int m = 0; m < ncomp; m++) P.p[m][t] = aos[t * ncomp + m];
}

__global__ void k_soa_to_aos(Ptrs P, int ncomp, long n, double* aos) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    for (int m = 0; m < ncomp; m++) aos[t * ncomp + m] = P.p[m][t];
}

__global__ void __launch_bounds__(256) k_nlevel(Level Lc, Level Lf, Fields F, int icrs, int jcrs, int kcrs) {
    nlevel_body(Lc, Lf, F, icrs, jcrs, kcrs, blockIdx.x, gridDim.x);
}
__global__ void __launch_bounds__(256) k_nlevel_z(ZoneGrid G, int n) {
    const ZoneBlock b = zone_block(G);
    const ZoneDesc& Z = G.z[b.z];
    nlevel_body(Z.lv[n], Z.lv[n - 1], Z.F, Z.icrs, Z.jcrs, Z.kcrs, b.bid, b.nblk);
}
__global__ void __launch_bounds__(256) k_prolong_fill(ProlongCtx P, Fields F) {
    prolong_fill_body(P, F, blockIdx.x, gridDim.x);
}
__global__ void __launch_bounds__(256) k_prolong_fill_z(ZoneGrid G, int n) {
    const ZoneBlock b = zone_block(G);
    const ZoneDesc& Z = G.z[b.z];
    const ProlongCtx P{Z.lv[n], Z.lv[n + 1], Z.icrs, Z.jcrs, Z.kcrs};
    prolong_fill_body(P, Z.F, b.bid, b.nblk);
}
template <int D, int R>
__global__ void __launch_bounds__(512) k_rs_sweep(Level L, Fields F, SegShape sh, double epn) {
    rs_sweep_body<D, R>(L, F, sh, epn, xcd_block(blockIdx.x, gridDim.x));
}
template <int D, int R>
__global__ void __launch_bounds__(512) k_rs_sweep_z(ZoneGrid G, int n, int P, int TI) {
    const ZoneBlock b = zone_block(G);
    const ZoneDesc& Z = G.z[b.z];
    rs_sweep_body<D, R>(Z.lv[n], Z.F, seg_lines(Z.lv[n], D, P, TI), Z.C.ep[1][D], xcd_block(b.bid, b.nblk));
}
// D = 2 also reads the pinned pressure into F.pfix (bcond_fm.F90:310-322),
// in the thread of the pinned node's k column (face node tfix) after that
// thread's own face writes: no other thread of the launch writes the
// column, and the i and j faces are complete
template <int D>
__global__ void __launch_bounds__(256) k_bc_fm(Level L, Fields F, Consts C, long tfix, long gfix) {
    bc_fm_body<D>(L, F, C, blockIdx.x);
    if (D == 2 && (long)blockIdx.x * blockDim.x + threadIdx.x == tfix) *F.pfix = F.q[0][gfix];
}
template <int D>
__global__ void __launch_bounds__(256) k_bc_fm_z(ZoneGrid G) {
    const ZoneBlock b = zone_block(G);
    const ZoneDesc& Z = G.z[b.z];
    bc_fm_body<D>(Z.lv[1], Z.F, Z.C, b.bid);
}
// one thread per zone
__global__ void k_pfix_read_z(const ZoneDesc* __restrict__ z, int nz) {
    const int t = (int)threadIdx.x;
    if (t < nz) *z[t].F.pfix = z[t].F.q[0][z[t].pfix];
}
__global__ void __launch_bounds__(256) k_pfix_sub(Level L, Fields F) { pfix_sub_body(L, F, blockIdx.x); }
__global__ void __launch_bounds__(256) k_pfix_sub_z(ZoneGrid G) {
    const ZoneBlock b = zone_block(G);
    pfix_sub_body(G.z[b.z].lv[1], G.z[b.z].F, b.bid);
}
__global__ void __launch_bounds__(256) k_resid_part(Level L, Fields F, int save, int pfix) {
    resid_part_body(L, F, save, pfix, blockIdx.x, gridDim.x);
}
__global__ void __launch_bounds__(256) k_resid_part_z(ZoneGrid G, int save) {
    const ZoneBlock b = zone_block(G);
    resid_part_body(G.z[b.z].lv[1], G.z[b.z].F, save, 0, b.bid, b.nblk);
}
// one workgroup per zone
__global__ void __launch_bounds__(256) k_resid_final_z(const ZoneDesc* __restrict__ z, int nblocks) {
    resid_final_body(z[blockIdx.x].F, nblocks);
}

inline unsigned nblk(long n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

void launch_inject(hipStream_t s, const Level* lv, int ng, int icrs, int jcrs, int kcrs, double* const* comps,
                   int ncomp) {
    Ptrs P{};
    for (int m = 0; m < ncomp && m < 16; m++) P.p[m] = comps[m];
    for (int n = 2; n <= ng; n++) {
        const Level& Lc = lv[n];
        const Level& Lf = lv[n - 1];
        hipLaunchKernelGGL(k_inject_p, dim3(nblk(Lc.n, 256)), dim3(256), 0, s, Lc, Lf, P, ncomp, icrs, jcrs, kcrs);
    }
}

void launch_nlevel(hipStream_t s, const Level& Lc, const Level& Lf, const Fields& F, int icrs, int jcrs, int kcrs) {
    hipLaunchKernelGGL(k_nlevel, dim3(nblk(Lc.n, 256)), dim3(256), 0, s, Lc, Lf, F, icrs, jcrs, kcrs);
}

void launch_zero_pk_boundary(hipStream_t s, const Level& L, const Fields& F) {
    hipLaunchKernelGGL(k_zero_pk_boundary, dim3(nblk(L.n, 256)), dim3(256), 0, s, L, F);
}

// the smoothing sweep of direction D over the interior lines of level L
template <int D>
static void rs_sweep(hipStream_t s, const Level& L, const Fields& F, double ep) {
    const int n = D == 0 ? L.ni - 2 : D == 1 ? L.nj - 2 : L.nk - 2;
    const long nl = D == 0 ? (long)(L.nj - 2) * (L.nk - 2) : D == 1 ? (long)(L.ni - 2) * (L.nk - 2)
                                                                    : (long)(L.ni - 2) * (L.nj - 2);
    // i lines: short segments (their rows are adjacent in memory, so lanes
    // reading row r of adjacent segments are R rows apart)
    const int R = (D == 0 || n <= 64) ? 4 : 8;
    const SegShape sh = seg_shape(n, nl, R, D == 0);
    const unsigned nb = (unsigned)((sh.nlines + sh.TI - 1) / sh.TI);
    const size_t lds = seg_lds_bytes(sh, 4);
    if (R == 4)
        hipLaunchKernelGGL((k_rs_sweep<D, 4>), dim3(nb), dim3(sh.TI * sh.P), lds, s, L, F, sh, ep);
    else
        hipLaunchKernelGGL((k_rs_sweep<D, 8>), dim3(nb), dim3(sh.TI * sh.P), lds, s, L, F, sh, ep);
}

void launch_prolong(hipStream_t s, const Level& Lf, const Level& Lc, const Fields& F, const Consts& C, int icrs,
                    int jcrs, int kcrs) {
    ProlongCtx P{Lf, Lc, icrs, jcrs, kcrs};
    hipLaunchKernelGGL(k_prolong_fill, dim3(nblk(fill_cells(P), 256)), dim3(256), 0, s, P, F);
    rs_sweep<0>(s, Lf, F, C.ep[1][0]);
    rs_sweep<1>(s, Lf, F, C.ep[1][1]);
    rs_sweep<2>(s, Lf, F, C.ep[1][2]);
}

void launch_outlet_ratios(hipStream_t s, const Level& L, const Fields& F, const int* btype, const int* bdir) {
    for (int b = 0; b < 4; b++) {
        if (btype[b] != DESALM_BC_INFLOW && btype[b] != DESALM_BC_OUTLET_MOC) continue;
        const long np = b < 2 ? (long)L.nj * L.nk : (long)L.ni * L.nk;
        hipLaunchKernelGGL(k_outlet_ratio, dim3(nblk(np, 256)), dim3(256), 0, s, L, F, b, bdir[b] - 1);
    }
}

void launch_bcond_fm(hipStream_t s, const Level& L, const Fields& F, const Consts& C, int ifix, int jfix, int kfix,
                     bool pfix_deferred) {
    const long tfix = (ifix - 1) + (long)L.ni * (jfix - 1);
    const long g = L.off + tfix + (long)L.ni * L.nj * (kfix - 1);
    hipLaunchKernelGGL(k_bc_fm<0>, dim3(nblk((long)L.nj * L.nk, 256)), dim3(256), 0, s, L, F, C, -1L, g);
    hipLaunchKernelGGL(k_bc_fm<1>, dim3(nblk((long)L.ni * L.nk, 256)), dim3(256), 0, s, L, F, C, -1L, g);
    hipLaunchKernelGGL(k_bc_fm<2>, dim3(nblk((long)L.ni * L.nj, 256)), dim3(256), 0, s, L, F, C, tfix, g);
    if (!pfix_deferred) hipLaunchKernelGGL(k_pfix_sub, dim3(nblk(L.n, 256)), dim3(256), 0, s, L, F);
}

// Partial sums of (q - qn)^2 over the fine interior (real_residual,
// src/main.F90:214-300), deterministic like k_resid_part; k_resid_final adds
// the parts into F.resid
__global__ void __launch_bounds__(256) k_real_part(Level L, Fields F) {
    __shared__ double sh[5][256];
    double acc[5] = {0, 0, 0, 0, 0};
    const int ni = L.ni, nj = L.nj, nk = L.nk;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < L.n; t += (long)gridDim.x * blockDim.x) {
        const int i = (int)(t % ni), j = (int)((t / ni) % nj), k = (int)(t / ((long)ni * nj));
        if (i == 0 || i == ni - 1 || j == 0 || j == nj - 1 || k == 0 || k == nk - 1) continue;
        const long g = L.off + t;
#pragma unroll
        for (int m = 0; m < 5; m++) {
            const double d = F.q[m][g] - F.qn[m][g];
            acc[m] += d * d;
        }
    }
#pragma unroll
    for (int m = 0; m < 5; m++) sh[m][threadIdx.x] = acc[m];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int m = 0; m < 5; m++) sh[m][threadIdx.x] += sh[m][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int m = 0; m < 5; m++) F.resid_part[blockIdx.x * 5 + m] = sh[m][0];
    }
}

// q(1:5) and xnut at n fine-level nodes (level-local index node[p]) into
// out[6 p .. 6 p + 5]
__global__ void k_probe(Fields F, int n, const long* node, double* out) {
    const int p = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (p >= n) return;
    const long g = node[p];
#pragma unroll
    for (int m = 0; m < 5; m++) out[6 * p + m] = F.q[m][g];
    out[6 * p + 5] = F.xnut[g];
}

void launch_real_residual(hipStream_t s, const Level& L, const Fields& F) {
    hipLaunchKernelGGL(k_real_part, dim3(kResidBlocks), dim3(256), 0, s, L, F);
    hipLaunchKernelGGL(k_resid_final, dim3(1), dim3(256), 0, s, F, kResidBlocks);
}

void launch_probe(hipStream_t s, const Fields& F, int n, const long* node, double* out) {
    if (n > 0) hipLaunchKernelGGL(k_probe, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, s, F, n, node, out);
}

void launch_resid_and_save(hipStream_t s, const Level& L, const Fields& F, int save, bool pfix) {
    hipLaunchKernelGGL(k_resid_part, dim3(kResidBlocks), dim3(256), 0, s, L, F, save, pfix ? 1 : 0);
    hipLaunchKernelGGL(k_resid_final, dim3(1), dim3(256), 0, s, F, kResidBlocks);
}

// ---- every zone of a group in one launch per kernel ----
void launch_nlevel_z(hipStream_t s, const ZoneList& Z, int n) {
    unsigned nb = 0;
    const ZoneGrid G = Z.grid([&](const ZoneDesc& z) { return nblk(z.lv[n].n, 256); }, nb);
    hipLaunchKernelGGL(k_nlevel_z, dim3(nb), dim3(256), 0, s, G, n);
}

template <int D>
static void rs_sweep_z(hipStream_t s, const ZoneList& Z, int n) {
    int len = 0;
    for (const ZoneDesc& z : Z.host) len = std::max(len, seg_lines(z.lv[n], D, 1, 1).n);
    const int R = (D == 0 || len <= 64) ? 4 : 8;
    const SegShape sh = seg_shape(len, 1, R, D == 0);
    unsigned nb = 0;
    const ZoneGrid G = Z.grid(
        [&](const ZoneDesc& z) { return (unsigned)((seg_lines(z.lv[n], D, sh.P, sh.TI).nlines + sh.TI - 1) / sh.TI); },
        nb);
    const size_t lds = seg_lds_bytes(sh, 4);
    if (R == 4)
        hipLaunchKernelGGL((k_rs_sweep_z<D, 4>), dim3(nb), dim3(sh.TI * sh.P), lds, s, G, n, sh.P, sh.TI);
    else
        hipLaunchKernelGGL((k_rs_sweep_z<D, 8>), dim3(nb), dim3(sh.TI * sh.P), lds, s, G, n, sh.P, sh.TI);
}

// prolongation from level n + 1 onto level n
void launch_prolong_z(hipStream_t s, const ZoneList& Z, int n) {
    unsigned nf = 0;
    const ZoneGrid Gf = Z.grid(
        [&](const ZoneDesc& z) { return nblk(fill_cells(ProlongCtx{z.lv[n], z.lv[n + 1], z.icrs, z.jcrs, z.kcrs}), 256); },
        nf);
    hipLaunchKernelGGL(k_prolong_fill_z, dim3(nf), dim3(256), 0, s, Gf, n);
    rs_sweep_z<0>(s, Z, n);
    rs_sweep_z<1>(s, Z, n);
    rs_sweep_z<2>(s, Z, n);
}

// bcond_fm faces and the pinned-pressure shift of every zone's fine level
void launch_bcond_fm_z(hipStream_t s, const ZoneList& Z) {
    unsigned fi = 0, fj = 0, fk = 0, gn = 0;
    const ZoneGrid Gi = Z.grid([&](const ZoneDesc& z) { return nblk((long)z.lv[1].nj * z.lv[1].nk, 256); }, fi);
    const ZoneGrid Gj = Z.grid([&](const ZoneDesc& z) { return nblk((long)z.lv[1].ni * z.lv[1].nk, 256); }, fj);
    const ZoneGrid Gk = Z.grid([&](const ZoneDesc& z) { return nblk((long)z.lv[1].ni * z.lv[1].nj, 256); }, fk);
    const ZoneGrid Gn = Z.grid([&](const ZoneDesc& z) { return nblk(z.lv[1].n, 256); }, gn);
    hipLaunchKernelGGL(k_bc_fm_z<0>, dim3(fi), dim3(256), 0, s, Gi);
    hipLaunchKernelGGL(k_bc_fm_z<1>, dim3(fj), dim3(256), 0, s, Gj);
    hipLaunchKernelGGL(k_bc_fm_z<2>, dim3(fk), dim3(256), 0, s, Gk);
    hipLaunchKernelGGL(k_pfix_read_z, dim3(1), dim3(64), 0, s, (const ZoneDesc*)Z.dev, Z.nz());
    hipLaunchKernelGGL(k_pfix_sub_z, dim3(gn), dim3(256), 0, s, Gn);
}

// update norms of every zone's fine level into its F.resid, then qold <- q
void launch_resid_and_save_z(hipStream_t s, const ZoneList& Z, int save) {
    unsigned nb = 0;
    const ZoneGrid G = Z.grid([&](const ZoneDesc&) { return (unsigned)kResidBlocks; }, nb);
    hipLaunchKernelGGL(k_resid_part_z, dim3(nb), dim3(256), 0, s, G, save);
    hipLaunchKernelGGL(k_resid_final_z, dim3(Z.nz()), dim3(256), 0, s, (const ZoneDesc*)Z.dev, kResidBlocks);
}

void launch_copy_qold(hipStream_t s, const Level& L, const Fields& F) {
    hipLaunchKernelGGL(k_copy_qold, dim3(nblk(L.n, 256)), dim3(256), 0, s, L, F);
}

void launch_aos_to_soa(hipStream_t s, const double* aos, int ncomp, long n, double* const* soa) {
    Ptrs P{};
    for (int m = 0; m < ncomp && m < 16; m++) P.p[m] = soa[m];
    hipLaunchKernelGGL(k_aos_to_soa, dim3(nblk(n, 256)), dim3(256), 0, s, aos, ncomp, n, P);
}

void launch_soa_to_aos(hipStream_t s, double* const* soa, int ncomp, long n, double* aos) {
    Ptrs P{};
    for (int m = 0; m < ncomp && m < 16; m++) P.p[m] = soa[m];
    hipLaunchKernelGGL(k_soa_to_aos, dim3(nblk(n, 256)), dim3(256), 0, s, P, ncomp, n, aos);
}

}  // namespace desalm
