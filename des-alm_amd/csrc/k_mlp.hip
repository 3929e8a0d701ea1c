// k_mlp.hip -- learned warm start of a physical time step.
//
// The reference starts every physical step's pseudo-iterations from the
// previous step's solution (mg_driver.F90:48-51: qn <- q, q unchanged).  The
// warm start replaces that initial guess on the fine level's interior by a
// two-step extrapolation in physical time whose coefficients are predicted
// per node and per variable,
//   q(0) = q^n + a d1 + b d2,  d1 = q^n - q^(n-1),  d2 = q^(n-1) - q^(n-2),
// with (a - 1, b) the outputs of a multilayer perceptron (zero outputs give
// linear extrapolation).  Its inputs are scale-free: for each of the five
// variables, the 18 time differences d1, d2 at the node and its eight
// neighbours (the six face neighbours and the second neighbours along i, the
// inflow direction; clamped to the block) divided by their root mean square,
// plus q^n at the node.  The
// pseudo-iterations then run to the same tolerance as from any other start,
// so the warm start only changes how many of them are needed.
//
// Everything is FP32 and written here: the feature gather, the dense layers
// (64 x 64 output tiles, the whole reduction dimension staged in LDS), the
// backward pass (data gradients with the same tiling, weight gradients as
// per-chunk partial products summed in a fixed order), the loss and Adam.
// Every layer's weights are one (K + 1) x N row-major matrix whose last row
// is the bias, and every activation matrix carries a trailing column of ones,
// so a layer is a single product.
#include "desalm_internal.h"

namespace desalm {

namespace {

constexpr int kBM = 64;     // rows (samples) per output tile
constexpr int kBN = 64;     // columns per output tile
constexpr int kKMax = 96;   // largest K + 1 a layer may have
constexpr int kWgRows = 512;  // rows per weight-gradient partial

// 64-bit mix of (seed, a, b) -> uniform index source (splitmix64 finaliser)
__device__ __forceinline__ unsigned long long mix3(unsigned long long s, unsigned long long a,
                                                   unsigned long long b) {
    unsigned long long z = s + 0x9E3779B97F4A7C15ull * (a + 1) + 0xBF58476D1CE4E5B9ull * (b + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// interior node (level-local index) of interior ordinal r, i fastest
__device__ __forceinline__ long interior_node(const Level& L, long r) {
    const long ni = L.ni - 2, nj = L.nj - 2;
    const long i = 1 + r % ni, j = 1 + (r / ni) % nj, k = 1 + r / (ni * nj);
    return i + (long)L.ni * (j + (long)L.nj * k);
}

// Features and (optionally) training target of level-local node t.
// Feature order: for component c, for stencil point s (node, -i, +i, -j,
// +j, -k, +k, -2i, +2i; the second neighbours clamped to i = 0 and ni - 1):
// d1, d2, each divided by the root mean square of the component's 18 (0
// when they are all zero); then q^n(c) at the node.
// mu / isd standardise the features when non-null.  Target: per component
// r = q - 2 q^n + q^(n-1), then d1 and d2 at the node (raw).
__device__ void ws_row(const Level& L, const Fields& F, double* const* qnm2, long t, float* x, float* y,
                       const float* mu, const float* isd) {
    const long g = L.off + t;
    const int i = (int)(t % L.ni);
    const long st[kWsPts] = {0, -1, 1, -(long)L.ni, (long)L.ni, -(long)L.ni * L.nj, (long)L.ni * L.nj,
                             i >= 2 ? -2 : -(long)i, i + 2 <= L.ni - 1 ? 2 : (long)(L.ni - 1 - i)};
    int f = 0;
    for (int c = 0; c < kWsComp; c++) {
        const double* a = F.qn[c];
        const double* b = F.qnm1[c];
        const double* d = qnm2[c];
        double v[2 * kWsPts];
        double ss = 0.0;
#pragma unroll
        for (int s = 0; s < kWsPts; s++) {
            const long p = g + st[s];
            const double bp = b[p];
            v[2 * s] = a[p] - bp;
            v[2 * s + 1] = bp - d[p - L.off];
            ss += v[2 * s] * v[2 * s] + v[2 * s + 1] * v[2 * s + 1];
        }
        const double rms = sqrt(ss / (2 * kWsPts));
#pragma unroll
        for (int e = 0; e < 2 * kWsPts; e++) {
            float w = rms > 0.0 ? (float)(v[e] / rms) : 0.0f;
            if (mu) w = (w - mu[f]) * isd[f];
            x[f++] = w;
        }
    }
    for (int c = 0; c < kWsComp; c++) {
        float w = (float)F.qn[c][g];
        if (mu) w = (w - mu[f]) * isd[f];
        x[f++] = w;
    }
    x[f] = 1.0f;  // bias column
    if (y) {
        for (int c = 0; c < kWsComp; c++) {
            const double qn = F.qn[c][g], qm = F.qnm1[c][g];
            y[c] = (float)(F.q[c][g] - 2.0 * qn + qm);
            y[kWsComp + c] = (float)(qn - qm);
            y[2 * kWsComp + c] = (float)(qm - qnm2[c][g - L.off]);
        }
    }
}

// record: m rows sampled from the interior with replacement (seeded),
// raw features into X (ld kWsIn + 1) and targets into T (ld kWsTgt)
__global__ void k_ws_sample(Level L, Fields F, WsQ Q, float* X, float* T, long m, unsigned long long seed,
                            long step) {
    const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= m) return;
    const long nint = (long)(L.ni - 2) * (L.nj - 2) * (L.nk - 2);
    const long t = interior_node(L, (long)(mix3(seed, (unsigned long long)step, (unsigned long long)r) % nint));
    ws_row(L, F, Q.q, t, X + r * (kWsIn + 1), T + r * kWsTgt, nullptr, nullptr);
}

// training batch: rows drawn from the first n rows of the sample buffer;
// features standardised, targets divided by the per-component scale
__global__ void k_ws_batch(const float* X, const float* T, long n, float* Xb, float* Tb, int m,
                           unsigned long long seed, long it, const float* stats) {
    const int r = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (r >= m) return;
    const long src = (long)(mix3(seed ^ 0x5851F42D4C957F2Dull, (unsigned long long)it, (unsigned long long)r) % n);
    const float* mu = stats;
    const float* isd = stats + kWsIn;
    const float* sc = stats + 2 * kWsIn;
    for (int f = 0; f < kWsIn; f++) Xb[(long)r * (kWsIn + 1) + f] = (X[src * (kWsIn + 1) + f] - mu[f]) * isd[f];
    Xb[(long)r * (kWsIn + 1) + kWsIn] = 1.0f;
    for (int e = 0; e < kWsTgt; e++) Tb[(long)r * kWsTgt + e] = T[src * kWsTgt + e] / sc[e % kWsComp];
}

// inference features of interior ordinals r0 .. r0 + m - 1, standardised
// when stats is not null
__global__ void k_ws_features(Level L, Fields F, WsQ Q, float* X, long r0, long m, const float* stats) {
    const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= m) return;
    ws_row(L, F, Q.q, interior_node(L, r0 + r), X + r * (kWsIn + 1), nullptr, stats,
           stats ? stats + kWsIn : nullptr);
}

// q(0) = q^n + (1 + y_a) d1 + y_b d2 on interior ordinals r0 .. r0 + m - 1
// (y = 0 without a network: linear extrapolation); nu~ keeps the floor of the
// SA update (rhs_sa_bcond.F90:283-287)
__global__ void k_ws_apply(Level L, Fields F, WsQ Q, const float* Y, long r0, long m) {
    const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= m) return;
    const long g = L.off + interior_node(L, r0 + r);
    for (int c = 0; c < kWsComp; c++) {
        const double qn = F.qn[c][g], qm = F.qnm1[c][g];
        double v;
        if (Y) {
            const double d1 = qn - qm, d2 = qm - Q.q[c][g - L.off];
            v = qn + (1.0 + (double)Y[r * kWsOut + c]) * d1 + (double)Y[r * kWsOut + kWsComp + c] * d2;
        } else {
            v = 2.0 * qn - qm;
        }
        if (c == 4) v = fmax(1.0e-07, v);
        F.q[c][g] = v;
    }
}

// Y[:, 0:N] = act(A[:, 0:K1] W), W (K1 x N) row-major; act 1 = tanh.
// With ones, Y[:, N] = 1.  256 threads, one 64 x 64 tile, 4 x 4 per thread.
__global__ void __launch_bounds__(256) k_dense(const float* A, int lda, int K1, const float* W, int N, float* Y,
                                               int ldy, int M, int act, int ones) {
    __shared__ float As[kKMax][kBM];
    __shared__ float Ws[kKMax][kBN];
    const int tid = (int)threadIdx.x;
    const int r0 = (int)blockIdx.x * kBM, c0 = (int)blockIdx.y * kBN;
    for (int e = tid; e < kBM * K1; e += 256) {
        const int row = e / K1, col = e % K1;
        As[col][row] = (r0 + row < M) ? A[(long)(r0 + row) * lda + col] : 0.0f;
    }
    for (int e = tid; e < K1 * kBN; e += 256) {
        const int k = e / kBN, n = e % kBN;
        Ws[k][n] = (c0 + n < N) ? W[(long)k * N + c0 + n] : 0.0f;
    }
    __syncthreads();
    const int tr = tid / 16, tc = tid % 16;
    float acc[4][4] = {};
    for (int k = 0; k < K1; k++) {
        float a[4], w[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            a[u] = As[k][tr * 4 + u];
            w[u] = Ws[k][tc * 4 + u];
        }
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
            for (int v = 0; v < 4; v++) acc[u][v] = fmaf(a[u], w[v], acc[u][v]);
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const int row = r0 + tr * 4 + u;
        if (row >= M) continue;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int col = c0 + tc * 4 + v;
            if (col < N) Y[(long)row * ldy + col] = act ? tanhf(acc[u][v]) : acc[u][v];
        }
        if (ones && blockIdx.y == 0 && tc == 0) Y[(long)row * ldy + N] = 1.0f;
    }
}

// dA[:, 0:K] = (dZ W[0:K, :]^T) * (1 - Aprev^2): the gradient through a
// layer (without its bias row) and the tanh that produced Aprev
__global__ void __launch_bounds__(256) k_dense_bwd(const float* dZ, int ldz, int N, const float* W, int K,
                                                   const float* Aprev, int lda, float* dA, int ldd, int M) {
    __shared__ float Zs[kBN][kBM];
    __shared__ float Wt[kBN][kBN];
    const int tid = (int)threadIdx.x;
    const int r0 = (int)blockIdx.x * kBM, c0 = (int)blockIdx.y * kBN;
    for (int e = tid; e < kBM * N; e += 256) {
        const int row = e / N, col = e % N;
        Zs[col][row] = (r0 + row < M) ? dZ[(long)(r0 + row) * ldz + col] : 0.0f;
    }
    for (int e = tid; e < N * kBN; e += 256) {
        const int n = e / kBN, k = e % kBN;
        Wt[n][k] = (c0 + k < K) ? W[(long)(c0 + k) * N + n] : 0.0f;
    }
    __syncthreads();
    const int tr = tid / 16, tc = tid % 16;
    float acc[4][4] = {};
    for (int n = 0; n < N; n++) {
        float a[4], w[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            a[u] = Zs[n][tr * 4 + u];
            w[u] = Wt[n][tc * 4 + u];
        }
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
            for (int v = 0; v < 4; v++) acc[u][v] = fmaf(a[u], w[v], acc[u][v]);
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const int row = r0 + tr * 4 + u;
        if (row >= M) continue;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int col = c0 + tc * 4 + v;
            if (col < K) {
                const float ap = Aprev[(long)row * lda + col];
                dA[(long)row * ldd + col] = acc[u][v] * (1.0f - ap * ap);
            }
        }
    }
}

// partial weight gradient of rows [g kWgRows, (g + 1) kWgRows):
// P[g][k][n] = sum_r A[r][k] dZ[r][n], k < K1 (the ones column gives the bias row)
__global__ void __launch_bounds__(256) k_wgrad_part(const float* A, int lda, int K1, const float* dZ, int ldz,
                                                    int N, int M, float* P) {
    __shared__ float As[kBM][kKMax];
    __shared__ float Zs[kBM][kBN];
    const int tid = (int)threadIdx.x;
    const int rbeg = (int)blockIdx.x * kWgRows;
    const int rend = min(M, rbeg + kWgRows);
    const int nout = K1 * N;
    constexpr int kPer = (kKMax * kBN + 255) / 256;
    float acc[kPer];
#pragma unroll
    for (int j = 0; j < kPer; j++) acc[j] = 0.0f;
    for (int rt = rbeg; rt < rend; rt += kBM) {
        const int nr = min(kBM, rend - rt);
        __syncthreads();
        for (int e = tid; e < kBM * K1; e += 256) {
            const int row = e / K1, col = e % K1;
            As[row][col] = row < nr ? A[(long)(rt + row) * lda + col] : 0.0f;
        }
        for (int e = tid; e < kBM * N; e += 256) {
            const int row = e / N, col = e % N;
            Zs[row][col] = row < nr ? dZ[(long)(rt + row) * ldz + col] : 0.0f;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kPer; j++) {
            const int o = tid + 256 * j;
            if (o < nout) {
                const int k = o / N, n = o % N;
                float s = acc[j];
                for (int r = 0; r < nr; r++) s = fmaf(As[r][k], Zs[r][n], s);
                acc[j] = s;
            }
        }
    }
#pragma unroll
    for (int j = 0; j < kPer; j++) {
        const int o = tid + 256 * j;
        if (o < nout) P[(long)blockIdx.x * nout + o] = acc[j];
    }
}

// G[o] = sum over the np partials, in order
__global__ void k_sum_parts(const float* P, int np, int nout, float* G) {
    const int o = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (o >= nout) return;
    float s = 0.0f;
    for (int g = 0; g < np; g++) s += P[(long)g * nout + o];
    G[o] = s;
}

// Loss of the coefficient form, in units of the per-component scale:
// e_c = y_a,c d1_c + y_b,c d2_c - r_c, loss = mean over rows and components
// of e^2; dZ = d loss / d y; per-block partial sums of e^2.
__global__ void __launch_bounds__(256) k_ws_loss(const float* Y, const float* T, int M, float* dZ, double* part) {
    __shared__ double sh[256];
    const long tot = (long)M * kWsComp;
    double s = 0.0;
    const float scale = 2.0f / (float)tot;
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < tot; e += (long)gridDim.x * 256) {
        const long r = e / kWsComp;
        const int c = (int)(e % kWsComp);
        const float* t = T + r * kWsTgt;
        const float* y = Y + r * kWsOut;
        const float d1 = t[kWsComp + c], d2 = t[2 * kWsComp + c];
        const float err = fmaf(y[c], d1, y[kWsComp + c] * d2) - t[c];
        dZ[r * kWsOut + c] = scale * err * d1;
        dZ[r * kWsOut + kWsComp + c] = scale * err * d2;
        s += (double)err * (double)err;
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void k_adam(float* w, const float* g, float* m, float* v, int n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
    const int o = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (o >= n) return;
    const float gi = g[o];
    const float mi = b1 * m[o] + (1.0f - b1) * gi;
    const float vi = b2 * v[o] + (1.0f - b2) * gi * gi;
    m[o] = mi;
    v[o] = vi;
    w[o] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// per-column sum and sum of squares over n rows (ld columns, first nc used),
// partial per block: part[blockIdx][2 nc]
__global__ void __launch_bounds__(256) k_colstats(const float* X, int ld, int nc, long n, double* part) {
    __shared__ double sh[256];
    const int c = (int)blockIdx.y;
    double s = 0.0, s2 = 0.0;
    for (long r = (long)blockIdx.x * 256 + threadIdx.x; r < n; r += (long)gridDim.x * 256) {
        const double v = X[r * ld + c];
        s += v;
        s2 += v * v;
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[((long)blockIdx.x * nc + c) * 2] = sh[0];
    __syncthreads();
    sh[threadIdx.x] = s2;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[((long)blockIdx.x * nc + c) * 2 + 1] = sh[0];
}

unsigned nblk(long n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

// ---- launchers ----

void launch_ws_sample(hipStream_t s, const Level& L, const Fields& F, const WsQ& Q, float* X, float* T, long m,
                      unsigned long long seed, long step) {
    hipLaunchKernelGGL(k_ws_sample, dim3(nblk(m, 128)), dim3(128), 0, s, L, F, Q, X, T, m, seed, step);
}

void launch_ws_batch(hipStream_t s, const float* X, const float* T, long n, float* Xb, float* Tb, int m,
                     unsigned long long seed, long it, const float* stats) {
    hipLaunchKernelGGL(k_ws_batch, dim3(nblk(m, 128)), dim3(128), 0, s, X, T, n, Xb, Tb, m, seed, it, stats);
}

void launch_ws_features(hipStream_t s, const Level& L, const Fields& F, const WsQ& Q, float* X, long r0, long m,
                        const float* stats) {
    hipLaunchKernelGGL(k_ws_features, dim3(nblk(m, 128)), dim3(128), 0, s, L, F, Q, X, r0, m, stats);
}

void launch_ws_apply(hipStream_t s, const Level& L, const Fields& F, const WsQ& Q, const float* Y, long r0, long m) {
    hipLaunchKernelGGL(k_ws_apply, dim3(nblk(m, 256)), dim3(256), 0, s, L, F, Q, Y, r0, m);
}

void launch_dense(hipStream_t s, const float* A, int lda, int K1, const float* W, int N, float* Y, int ldy, int M,
                  int act, int ones) {
    hipLaunchKernelGGL(k_dense, dim3(nblk(M, kBM), nblk(N, kBN)), dim3(256), 0, s, A, lda, K1, W, N, Y, ldy, M, act,
                       ones);
}

void launch_dense_bwd(hipStream_t s, const float* dZ, int ldz, int N, const float* W, int K, const float* Aprev,
                      int lda, float* dA, int ldd, int M) {
    hipLaunchKernelGGL(k_dense_bwd, dim3(nblk(M, kBM), nblk(K, kBN)), dim3(256), 0, s, dZ, ldz, N, W, K, Aprev, lda,
                       dA, ldd, M);
}

int wgrad_parts(int M) { return (int)nblk(M, kWgRows); }

void launch_wgrad(hipStream_t s, const float* A, int lda, int K1, const float* dZ, int ldz, int N, int M, float* P,
                  float* G) {
    const int np = wgrad_parts(M);
    hipLaunchKernelGGL(k_wgrad_part, dim3(np), dim3(256), 0, s, A, lda, K1, dZ, ldz, N, M, P);
    hipLaunchKernelGGL(k_sum_parts, dim3(nblk((long)K1 * N, 256)), dim3(256), 0, s, P, np, K1 * N, G);
}

void launch_ws_loss(hipStream_t s, const float* Y, const float* T, int M, float* dZ, double* part, int nparts) {
    hipLaunchKernelGGL(k_ws_loss, dim3(nparts), dim3(256), 0, s, Y, T, M, dZ, part);
}

void launch_adam(hipStream_t s, float* w, const float* g, float* m, float* v, int n, float lr, float b1, float b2,
                 float eps, float c1, float c2) {
    hipLaunchKernelGGL(k_adam, dim3(nblk(n, 256)), dim3(256), 0, s, w, g, m, v, n, lr, b1, b2, eps, c1, c2);
}

void launch_colstats(hipStream_t s, const float* X, int ld, int nc, long n, double* part, int nparts) {
    hipLaunchKernelGGL(k_colstats, dim3(nparts, nc), dim3(256), 0, s, X, ld, nc, n, part);
}

}  // namespace desalm
