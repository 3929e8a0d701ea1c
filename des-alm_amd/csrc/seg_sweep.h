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

// LDS values per segment slot: the larger of 4 (a Moebius map) and 2 NS
__host__ __device__ constexpr int seg_slot(int ns) { return 2 * ns > 4 ? 2 * ns : 4; }

// Segments of R rows; j and k lines (seg_fast = false) put 16 adjacent lines
// (adjacent i) on adjacent lanes, i lines (seg_fast = true) put the segments
// of one line on adjacent lanes.
inline SegShape seg_shape(int n, long nlines, int R, bool seg_fast) {
    SegShape sh;
    sh.n = n;
    sh.nlines = nlines;
    sh.P = (n + R - 1) / R;
    if (sh.P < 4) sh.P = 4;
    sh.TI = seg_fast ? (256 / sh.P > 1 ? 256 / sh.P : 1) : 16;
    while (sh.TI * sh.P > 512 && sh.TI > 1) sh.TI /= 2;
    return sh;
}

// The solve.  On entry a, b, c, d hold the thread's nr rows (row r is line
// row s0 + r), a_next the subdiagonal of row s0 + nr when that row exists
// (for the interchange test).  On exit d holds the solution x.  Every thread
// of the workgroup must call it (it contains barriers); lds holds
// TI * P * seg_slot(NS) values of S; slot = this thread's slot index
// li * P + ps.
template <int R, int NS, typename S>
__device__ void seg_solve(S (&a)[R], S (&b)[R], S (&c)[R], S (&d)[R][NS], S a_next, int nr, int n, int s0,
                          S* lds, int P, int TI, long nlines, long line0, int li, int ps, unsigned& piv,
                          int& zrow) {
    constexpr int K = seg_slot(NS);
    const int tid = (int)threadIdx.x;
    S* slot = lds + (long)(li * P + ps) * K;
    const int sl = tid / NS, sm = tid % NS;  // scanner: line, right-hand side
    const bool scanner = tid < NS * TI && line0 + sl < nlines;

    // phase A: Moebius maps of c' over the segment
    {
        S m0 = S(1), m1 = S(0), m2 = S(0), m3 = S(1);
#pragma unroll
        for (int r = 0; r < R; r++) {
            if (r < nr) {
                // [c'; 1] -> [c; b - a c']
                const S n0 = c[r] * m2, n1 = c[r] * m3;
                const S n2 = b[r] * m2 - a[r] * m0, n3 = b[r] * m3 - a[r] * m1;
                const S sc = fmax(fmax(fabs(n0), fabs(n1)), fmax(fabs(n2), fabs(n3)));
                const S is = sc > S(0) ? S(1) / sc : S(1);
                m0 = n0 * is; m1 = n1 * is; m2 = n2 * is; m3 = n3 * is;
            }
        }
        slot[0] = m0; slot[1] = m1; slot[2] = m2; slot[3] = m3;
    }
    __syncthreads();
    if (scanner && sm == 0) {
        S cin = S(0);
        for (int sg = 0; sg < P; sg++) {
            S* q = lds + (long)(sl * P + sg) * K;
            const S q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
            q[0] = cin;
            cin = (q0 * cin + q1) / (q2 * cin + q3);
        }
    }
    __syncthreads();
    const S cst = slot[0];

    // phase B: exact c' over the rows, the interchange test, and affine maps
    // of d' over the segment
    S cp[R];
    {
        S U[NS], V[NS];
#pragma unroll
        for (int m = 0; m < NS; m++) { U[m] = S(0); V[m] = S(1); }
        S cprev = cst;
#pragma unroll
        for (int r = 0; r < R; r++) {
            if (r < nr) {
                const S den = b[r] - a[r] * cprev;
                const S inv = S(1) / den;
                cprev = c[r] * inv;
                cp[r] = cprev;
                if (!(den != S(0) && isfinite(den)) && zrow == 0) zrow = s0 + r + 1;
                if (s0 + r + 1 < n && fabs(den) < fabs(r + 1 < nr ? a[r + 1 < R ? r + 1 : R - 1] : a_next)) piv++;
                const S v = -a[r] * inv;
#pragma unroll
                for (int m = 0; m < NS; m++) {
                    U[m] = d[r][m] * inv + v * U[m];
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
        for (int sg = 0; sg < P; sg++) {
            S* q = lds + (long)(sl * P + sg) * K + 2 * sm;
            const S u = q[0], v = q[1];
            q[0] = din;
            din = u + v * din;
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
                const S den = b[r] - a[r] * (r == 0 ? cst : cp[r > 0 ? r - 1 : 0]);
                const S inv = S(1) / den;
#pragma unroll
                for (int m = 0; m < NS; m++) {
                    dprev[m] = (d[r][m] - a[r] * dprev[m]) * inv;
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
#pragma unroll
                for (int m = 0; m < NS; m++) {
                    A[m] = d[r][m] - cp[r] * A[m];
                    B[m] = -cp[r] * B[m];
                }
            }
        }
#pragma unroll
        for (int m = 0; m < NS; m++) { slot[2 * m] = A[m]; slot[2 * m + 1] = B[m]; }
    }
    __syncthreads();
    if (scanner) {
        S xin = S(0);
        for (int sg = P - 1; sg >= 0; sg--) {
            S* q = lds + (long)(sl * P + sg) * K + 2 * sm;
            const S aa = q[0], bb = q[1];
            q[0] = xin;
            xin = aa + bb * xin;
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
#pragma unroll
            for (int m = 0; m < NS; m++) {
                x[m] = d[r][m] - cp[r] * x[m];
                d[r][m] = x[m];
            }
        }
    }
}

}  // namespace desalm
