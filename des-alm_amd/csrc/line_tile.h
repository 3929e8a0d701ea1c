// line_tile.h -- LDS staging of i-direction line segments.
//
// Interior i-lines are contiguous in memory, so a thread per line that walks
// its own line touches a different cache line per lane on every row.  The
// i-direction line solvers instead let a workgroup own kTL lines and move
// chunks of kTR consecutive rows of all of them between HBM and LDS with
// row-fastest (coalesced) accesses; each thread then runs its recurrence out
// of LDS.  The arithmetic of the recurrences is untouched.
#pragma once

#include "desalm_internal.h"

namespace desalm {

constexpr int kTL = 32;       // i-lines per workgroup
constexpr int kTR = 16;       // rows per chunk
constexpr int kTS = kTR + 1;  // LDS stride of one line's chunk (odd: no bank conflicts)

// level-local node index of i = 0 on interior i-line t = (j-1) + (nj-2)(k-1)
__device__ __forceinline__ long iline_base(const Level& L, long t) {
    const int nj2 = L.nj - 2;
    return (long)L.ni * ((1 + t % nj2) + (long)L.nj * (1 + t / nj2));
}

// Copies between HBM and an LDS tile of LINES lines x NRW rows (LDS stride
// STRIDE) of element type T.  Row r of line l lives at g[lb[l] + r0 + r];
// rows with r0 + r >= rmax and lines l >= nl are skipped (loads give 0).  NT
// threads take the elements row-fastest.
template <int NT, int LINES, int NRW, int STRIDE, typename T = double>
struct TileIO {
    static constexpr int kN = LINES * NRW;
    static constexpr int kPer = (kN + NT - 1) / NT;

    __device__ static void load(T (&v)[kPer], const T* __restrict__ g, const long* lb, int nl, int r0, int rmax) {
#pragma unroll
        for (int it = 0; it < kPer; it++) {
            const int e = (int)threadIdx.x + NT * it;
            const int l = e / NRW, r = e % NRW;
            const bool ok = e < kN && l < nl && r0 + r < rmax;
            v[it] = ok ? g[lb[l] + r0 + r] : T(0);
        }
    }
    __device__ static void to_lds(const T (&v)[kPer], T* lds) {
#pragma unroll
        for (int it = 0; it < kPer; it++) {
            const int e = (int)threadIdx.x + NT * it;
            if (e < kN) lds[(e / NRW) * STRIDE + e % NRW] = v[it];
        }
    }
    __device__ static void store(const T* lds, T* __restrict__ g, const long* lb, int nl, int r0, int rmax) {
#pragma unroll
        for (int it = 0; it < kPer; it++) {
            const int e = (int)threadIdx.x + NT * it;
            const int l = e / NRW, r = e % NRW;
            if (e < kN && l < nl && r0 + r < rmax) g[lb[l] + r0 + r] = lds[l * STRIDE + r];
        }
    }
};

}  // namespace desalm
