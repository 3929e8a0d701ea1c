"""CPU restatement of the learned warm start (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's CPU baseline may use this
file, as the checker of the HIP path; the library never calls it.

The reference has no learned part: it starts every physical step from the
previous solution (src/mg_driver.F90:48-51).  What is restated here is the
library's own definition (include/desalm.h, desalm_ws_*):
  features  per interior node, for c = p, u, v, w, nu~: the 18 differences
            d1 = q^n - q^(n-1), d2 = q^(n-1) - q^(n-2) at the stencil (node,
            -i, +i, -j, +j, -k, +k, -2i, +2i; the second neighbours clamped
            to i = 0 and ni - 1), divided by their root mean square (0 when
            all are 0); then q^n at the node (95 values)
  targets   per c: r = q^(n+1) - 2 q^n + q^(n-1), d1, d2 at the node
  network   95 -> tanh(nh) -> tanh(nh) -> 10, each layer one (K+1) x N matrix
            whose last row is the bias; outputs (a - 1, b) per variable
  loss      mean over rows and variables of ((a - 1) d1 + b d2 - r)^2, all
            three in units of the variable's scale
  start     q^n + a d1 + b d2 on the interior, nu~ >= 1e-7
Everything is float64 here.  Pinning: loss_grad against central finite
differences of loss (tests/test_mlp_oracle.py), forward against a direct
per-sample evaluation, features against hand-indexed values, and the zero
network against plain linear extrapolation.
"""
import numpy as np

N_IN, N_OUT, N_COMP = 95, 10, 5
OFFS = ((0, 0, 0), (-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1), (-2, 0, 0), (2, 0, 0))


def param_count(nh):
    return (N_IN + 1) * nh + (nh + 1) * nh + (nh + 1) * N_OUT


def unpack(w, nh):
    w = np.asarray(w, dtype=np.float64)
    o2 = (N_IN + 1) * nh
    o3 = o2 + (nh + 1) * nh
    return (w[:o2].reshape(N_IN + 1, nh), w[o2:o3].reshape(nh + 1, nh), w[o3:].reshape(nh + 1, N_OUT))


def interior_index(dims):
    """level-local node index of each interior ordinal (i fastest)"""
    ni, nj, nk = dims
    i = np.arange(1, ni - 1)
    j = np.arange(1, nj - 1)
    k = np.arange(1, nk - 1)
    K, J, I = np.meshgrid(k, j, i, indexing="ij")
    return (I + ni * (J + nj * K)).ravel()


def features(qn, qnm1, qnm2, dims):
    """(m, 95) float32 features of every interior node, computed as the
    library does (float64 differences and root mean square, one rounding)"""
    ni, nj, _ = dims
    g = interior_index(dims)
    gi = g % ni
    cols = []
    for c in range(N_COMP):
        v = []
        for di, dj, dk in OFFS:
            if abs(di) == 2:  # second neighbours along i, clamped to the block
                di = np.clip(gi + di, 0, ni - 1) - gi
            p = g + di + ni * (dj + nj * dk)
            v.append(qn[p, c] - qnm1[p, c])
            v.append(qnm1[p, c] - qnm2[p, c])
        v = np.stack(v, axis=1)
        ss = np.zeros(len(g))
        for s in range(len(OFFS)):
            ss = ss + (v[:, 2 * s] * v[:, 2 * s] + v[:, 2 * s + 1] * v[:, 2 * s + 1])
        rms = np.sqrt(ss / (2.0 * len(OFFS)))
        safe = np.where(rms > 0.0, rms, 1.0)
        cols.append(np.where(rms[:, None] > 0.0, v / safe[:, None], 0.0).astype(np.float32))
    cols.append(qn[g].astype(np.float32))
    return np.concatenate(cols, axis=1)


def targets(q, qn, qnm1, qnm2, dims):
    """(m, 15): r, d1, d2 per variable (float64)"""
    g = interior_index(dims)
    return np.concatenate([q[g] - 2.0 * qn[g] + qnm1[g], qn[g] - qnm1[g], qnm1[g] - qnm2[g]], axis=1)


def forward(w, x, nh, keep=False):
    W1, W2, W3 = unpack(w, nh)
    a0 = np.hstack([x, np.ones((x.shape[0], 1))])
    a1 = np.tanh(a0 @ W1)
    a1e = np.hstack([a1, np.ones((x.shape[0], 1))])
    a2 = np.tanh(a1e @ W2)
    a2e = np.hstack([a2, np.ones((x.shape[0], 1))])
    y = a2e @ W3
    return (y, (a0, a1e, a2e)) if keep else y


def residual(y, t):
    """(a - 1) d1 + b d2 - r per row and variable"""
    return y[:, :N_COMP] * t[:, N_COMP:2 * N_COMP] + y[:, N_COMP:] * t[:, 2 * N_COMP:] - t[:, :N_COMP]


def loss(w, x, t, nh):
    return float(np.mean(residual(forward(w, x, nh), t) ** 2))


def loss_grad(w, x, t, nh):
    """(loss, gradient with respect to the packed parameters)"""
    W1, W2, W3 = unpack(w, nh)
    y, (a0, a1e, a2e) = forward(w, x, nh, keep=True)
    m = x.shape[0]
    e = residual(y, t)
    s = 2.0 / (m * N_COMP)
    d3 = np.hstack([s * e * t[:, N_COMP:2 * N_COMP], s * e * t[:, 2 * N_COMP:]])
    g3 = a2e.T @ d3
    d2 = (d3 @ W3[:nh].T) * (1.0 - a2e[:, :nh] ** 2)
    g2 = a1e.T @ d2
    d1 = (d2 @ W2[:nh].T) * (1.0 - a1e[:, :nh] ** 2)
    g1 = a0.T @ d1
    return float(np.mean(e ** 2)), np.concatenate([g1.ravel(), g2.ravel(), g3.ravel()])


def standardise(x, stats):
    mu, isd = stats[:N_IN], stats[N_IN:2 * N_IN]
    return (x - mu) * isd


def warm_start(q, qn, qnm1, qnm2, dims, w=None, nh=None, stats=None):
    """the fine level's q after the warm start: linear extrapolation, or the
    network's coefficients when w is given"""
    out = q.copy()
    g = interior_index(dims)
    d1 = qn[g] - qnm1[g]
    if w is None:
        v = 2.0 * qn[g] - qnm1[g]
    else:
        x = standardise(features(qn, qnm1, qnm2, dims).astype(np.float64), stats)
        y = forward(w, x, nh)
        v = qn[g] + (1.0 + y[:, :N_COMP]) * d1 + y[:, N_COMP:] * (qnm1[g] - qnm2[g])
    v[:, 4] = np.maximum(1.0e-7, v[:, 4])
    out[g] = v
    return out
