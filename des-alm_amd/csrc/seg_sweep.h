// seg_sweep.h -- tridiagonal line solves with every row on chip.
//
// A workgroup owns TI lines and cuts each into P segments of at most R
// consecutive rows; the thread of a segment holds its rows' coefficients and
// right-hand sides in registers.  Thomas elimination is three first-order
// recurrences:
//   c'_s = c_s / (b_s - a_s c'_(s-1))            (a Moebius map of c'_(s-1))
//   d'_s = (d_s - a_s d'_(s-1)) / den_s          (affine in d'_(s-1))
//   x_s  = d'_s - c'_s x_(s+1)                   (affine in x_(s+1))
// Each thread composes its rows' maps; one thread per (line, right-hand side)
// walks the P segment maps in LDS and hands every segment the value entering
// it; every thread then runs the exact sequential recurrence over its own rows
// from that value.  Rows are therefore eliminated as LAPACK SGTSV eliminates
// them when it does not interchange rows, and the values at segment starts
// differ from a one-thread sweep only by rounding.  SGTSV's interchange test
// (|den_s| < |a_(s+1)|) is evaluated on every row and counted; the first row
// whose pivot is zero or not finite is SGTSV's INFO.
//
// Used by the SA ADI sweeps and the prolongation smoothing sweeps; the DAF
// sweeps (k_daf.hip) follow the same scheme with three matrices per line.
#pragma once

#include "desalm_internal.h"

namespace desalm {

// launch shape of one segmented sweep
struct SegShape {
    int n;        // rows per line
    int P;        // segments per line (>= 4: one scanning thread per right-hand side)
    int TI;       // lines per workgroup
    long nlines;
};

// LDS values per segment slot: the larger of 4 (a Moebius map) and 2 NS,
// rounded up to an odd count, and per line its P slots rounded up to an odd
// count: the threads of adjacent segments, and the scanning threads of
// adjacent lines, then fall on different LDS banks
__host__ __device__ constexpr int seg_slot(int ns) { return (2 * ns > 4 ? 2 * ns : 4) | 1; }
__host__ __device__ constexpr int seg_line_stride(int P, int ns) { return (P * seg_slot(ns)) | 1; }
inline size_t seg_lds_bytes(const SegShape& sh, int ns) {
    return sizeof(double) * (size_t)sh.TI * seg_line_stride(sh.P, ns);
}

// Segments of R rows; j and k lines (seg_fast = false) put 16 adjacent lines
// (adjacent i) on adjacent lanes, i lines (seg_fast = true) put the segments
// of one line on adjacent lanes.
inline SegShape seg_shape(int n, long nlines, int R, bool seg_fast) {
    SegShape sh;
    sh.n = n;
    sh.nlines = nlines;
    sh.P = (n + R - 1) / R;
    if (sh.P < 4) sh.P = 4;
    // j and k lines: 16 adjacent i lines per workgroup (128-byte rows; the
    // launches order workgroups XCD-contiguously, so neighbouring runs share
    // their edge cache lines in one L2)
    sh.TI = seg_fast ? (256 / sh.P > 1 ? 256 / sh.P : 1) : 16;
    while (sh.TI * sh.P > 512 && sh.TI > 1) sh.TI /= 2;
    return sh;
}

// lines of direction D over the interior of level L with the launch's P, TI
inline __host__ __device__ SegShape seg_lines(const Level& L, int D, int P, int TI) {
    SegShape sh;
    sh.n = D == 0 ? L.ni - 2 : D == 1 ? L.nj - 2 : L.nk - 2;
    sh.nlines = D == 0 ? (long)(L.nj - 2) * (L.nk - 2) : D == 1 ? (long)(L.ni - 2) * (L.nk - 2)
                                                                : (long)(L.ni - 2) * (L.nj - 2);
    sh.P = P;
    sh.TI = TI;
    return sh;
}

// exponent e of mx = f 2^e with f in [0.5, 1): scaling by 2^-e is exact
__device__ __forceinline__ int seg_exponent(double mx) {
    int e = 0;
    (void)frexp(mx, &e);
    return e;
}

// The solve.  On entry a, b, c, d hold the thread's nr rows (row r is line
// row s0 + r), a_next the subdiagonal of row s0 + nr when that row exists
// (for the interchange test).  On exit d holds the solution x.  Every thread
// of the workgroup must call it (it contains barriers); lds holds
// TI * seg_line_stride(P, NS) values of S; this thread's slot is segment ps
// of line li.
template <int R, int NS, typename S>
__device__ void seg_solve(S (&a)[R], S (&b)[R], S (&c)[R], S (&d)[R][NS], S a_next, int nr, int n, int s0,
                          S* lds, int P, int TI, long nlines, long line0, int li, int ps, unsigned& piv,
                          int& zrow) {
    constexpr int K = seg_slot(NS);
    const int LS = seg_line_stride(P, NS);
    const int tid = (int)threadIdx.x;
    S* slot = lds + li * LS + ps * K;
    const int sl = tid / NS, sm = tid % NS;  // scanner: line, right-hand side
    const bool scanner = tid < NS * TI && line0 + sl < nlines;
    S* sline = lds + sl * LS;

    // phase A: Moebius maps of c' over the segment
    {
        S m0 = S(1), m1 = S(0), m2 = S(0), m3 = S(1);
#pragma unroll
        for (int r = 0; r < R; r++) {
            if (r < nr) {
                // [c'; 1] -> [c; b - a c']
                const S n0 = c[r] * m2, n1 = c[r] * m3;
                const S n2 = b[r] * m2 - a[r] * m0, n3 = b[r] * m3 - a[r] * m1;
                const int e = -seg_exponent(fmax(fmax(fabs(n0), fabs(n1)), fmax(fabs(n2), fabs(n3))));
                m0 = ldexp(n0, e); m1 = ldexp(n1, e); m2 = ldexp(n2, e); m3 = ldexp(n3, e);
            }
        }
        slot[0] = m0; slot[1] = m1; slot[2] = m2; slot[3] = m3;
    }
    __syncthreads();
    if (scanner && sm == 0) {
        // projective pair (cn, cd): c' entering the segment is cn / cd
        S cn = S(0), cd = S(1);
        for (int sg0 = 0; sg0 < P; sg0 += 4) {
            S q[4][4];
#pragma unroll
            for (int u = 0; u < 4; u++)
                if (sg0 + u < P)
#pragma unroll
                    for (int e = 0; e < 4; e++) q[u][e] = sline[(sg0 + u) * K + e];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                if (sg0 + u < P) {
                    sline[(sg0 + u) * K] = cn / cd;
                    const S nn = q[u][0] * cn + q[u][1] * cd, dd = q[u][2] * cn + q[u][3] * cd;
                    const int e = -seg_exponent(fmax(fabs(nn), fabs(dd)));
                    cn = ldexp(nn, e);
                    cd = ldexp(dd, e);
                }
            }
        }
    }
    __syncthreads();
    const S cst = slot[0];

    // phase B: exact c' over the rows (their 1 / den kept), the interchange
    // test, and affine maps of d' over the segment
    S inv[R];
    {
        S U[NS], V[NS];
#pragma unroll
        for (int m = 0; m < NS; m++) { U[m] = S(0); V[m] = S(1); }
        S cprev = cst;
#pragma unroll
        for (int r = 0; r < R; r++) {
            inv[r] = S(0);
            if (r < nr) {
                const S den = b[r] - a[r] * cprev;
                const S iv = S(1) / den;
                inv[r] = iv;
                cprev = c[r] * iv;
                if (!(den != S(0) && isfinite(den)) && zrow == 0) zrow = s0 + r + 1;
                if (s0 + r + 1 < n && fabs(den) < fabs(r + 1 < nr ? a[r + 1 < R ? r + 1 : R - 1] : a_next)) piv++;
                const S v = -a[r] * iv;
#pragma unroll
                for (int m = 0; m < NS; m++) {
                    U[m] = d[r][m] * iv + v * U[m];
                    V[m] = v * V[m];
                }
            }
        }
        // between two scans a slot is read and written by its own thread only
#pragma unroll
        for (int m = 0; m < NS; m++) { slot[2 * m] = U[m]; slot[2 * m + 1] = V[m]; }
    }
    __syncthreads();
    if (scanner) {
        S din = S(0);
        for (int sg0 = 0; sg0 < P; sg0 += 8) {
            S u[8], v[8];
#pragma unroll
            for (int k = 0; k < 8; k++)
                if (sg0 + k < P) { u[k] = sline[(sg0 + k) * K + 2 * sm]; v[k] = sline[(sg0 + k) * K + 2 * sm + 1]; }
#pragma unroll
            for (int k = 0; k < 8; k++) {
                if (sg0 + k < P) {
                    sline[(sg0 + k) * K + 2 * sm] = din;
                    din = u[k] + v[k] * din;
                }
            }
        }
    }
    __syncthreads();

    // phase C: exact d' over the rows; affine maps of x, scanned backwards
    {
        S dprev[NS];
#pragma unroll
        for (int m = 0; m < NS; m++) dprev[m] = slot[2 * m];
#pragma unroll
        for (int r = 0; r < R; r++) {
            if (r < nr) {
#pragma unroll
                for (int m = 0; m < NS; m++) {
                    dprev[m] = (d[r][m] - a[r] * dprev[m]) * inv[r];
                    d[r][m] = dprev[m];
                }
            }
        }
        S A[NS], B[NS];
#pragma unroll
        for (int m = 0; m < NS; m++) { A[m] = S(0); B[m] = S(1); }
#pragma unroll
        for (int r = R - 1; r >= 0; r--) {
            if (r < nr) {
                const S cp = c[r] * inv[r];
#pragma unroll
                for (int m = 0; m < NS; m++) {
                    A[m] = d[r][m] - cp * A[m];
                    B[m] = -cp * B[m];
                }
            }
        }
#pragma unroll
        for (int m = 0; m < NS; m++) { slot[2 * m] = A[m]; slot[2 * m + 1] = B[m]; }
    }
    __syncthreads();
    if (scanner) {
        S xin = S(0);
        for (int sg1 = P - 1; sg1 >= 0; sg1 -= 8) {
            S u[8], v[8];
#pragma unroll
            for (int k = 0; k < 8; k++)
                if (sg1 - k >= 0) { u[k] = sline[(sg1 - k) * K + 2 * sm]; v[k] = sline[(sg1 - k) * K + 2 * sm + 1]; }
#pragma unroll
            for (int k = 0; k < 8; k++) {
                if (sg1 - k >= 0) {
                    sline[(sg1 - k) * K + 2 * sm] = xin;
                    xin = u[k] + v[k] * xin;
                }
            }
        }
    }
    __syncthreads();

    // phase D: back substitution over the rows
    S x[NS];
#pragma unroll
    for (int m = 0; m < NS; m++) x[m] = slot[2 * m];
#pragma unroll
    for (int r = R - 1; r >= 0; r--) {
        if (r < nr) {
            const S cp = c[r] * inv[r];
#pragma unroll
            for (int m = 0; m < NS; m++) {
                x[m] = d[r][m] - cp * x[m];
                d[r][m] = x[m];
            }
        }
    }
}

}  // namespace desalm
