"""Pins the CPU restatement of the learned warm start (oracle/mlp_oracle.py):
the analytic gradient against central finite differences, the batched
forward pass against a per-sample evaluation, the feature gather against
hand-indexed values, and the zero network against linear extrapolation."""
import numpy as np

import mlp_oracle as mo


def rand_params(nh, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(0.0, 0.3, mo.param_count(nh))


def rand_rows(m, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(m, mo.N_IN)), rng.normal(size=(m, 3 * mo.N_COMP))


def test_gradient_matches_finite_differences():
    nh = 6
    w = rand_params(nh, 2)
    x, t = rand_rows(17, 1)
    loss, g = mo.loss_grad(w, x, t, nh)
    assert abs(loss - mo.loss(w, x, t, nh)) < 1e-14
    # a spread of weights and every bias of each of the three layers
    o2 = (mo.N_IN + 1) * nh
    o3 = o2 + (nh + 1) * nh
    idx = list(range(0, o2, 37)) + list(range(o2 - nh, o2)) + list(range(o2, o3, 5)) + list(range(o3, len(w)))
    h = 1e-6
    for i in idx:
        wp = w.copy()
        wm = w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (mo.loss(wp, x, t, nh) - mo.loss(wm, x, t, nh)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 * max(1.0, abs(g[i])), i


def test_gradient_catches_swapped_coefficients():
    # a plausible slip (a paired with d2 and b with d1) must fail the check
    nh = 5
    w = rand_params(nh, 4)
    x, t = rand_rows(9, 3)
    _, g = mo.loss_grad(w, x, t, nh)
    t_sw = t.copy()
    t_sw[:, 5:10], t_sw[:, 10:15] = t[:, 10:15], t[:, 5:10]
    _, g_sw = mo.loss_grad(w, x, t_sw, nh)
    assert np.abs(g - g_sw).max() > 1e-6


def test_forward_matches_per_sample_evaluation():
    nh = 7
    w = rand_params(nh, 6)
    x, _ = rand_rows(5, 5)
    W1, W2, W3 = mo.unpack(w, nh)
    y = mo.forward(w, x, nh)
    for r in range(5):
        h1 = np.tanh(W1[:-1].T @ x[r] + W1[-1])
        h2 = np.tanh(W2[:-1].T @ h1 + W2[-1])
        assert np.allclose(W3[:-1].T @ h2 + W3[-1], y[r], rtol=0, atol=1e-13)


def test_features_hand_indexed():
    dims = (6, 5, 4)
    n1 = 6 * 5 * 4
    rng = np.random.default_rng(7)
    qn, qm1, qm2 = (rng.normal(size=(n1, 5)) for _ in range(3))
    x = mo.features(qn, qm1, qm2, dims)
    assert x.shape == (4 * 3 * 2, mo.N_IN)
    node = lambda a, b, c: a + 6 * (b + 5 * c)  # noqa: E731

    def stencil(i, j, k):
        # node, -i, +i, -j, +j, -k, +k, then the second neighbours along i,
        # clamped to the block (i = 0 .. 5)
        return [node(i, j, k), node(i - 1, j, k), node(i + 1, j, k), node(i, j - 1, k), node(i, j + 1, k),
                node(i, j, k - 1), node(i, j, k + 1), node(max(i - 2, 0), j, k), node(min(i + 2, 5), j, k)]

    # interior ordinal 5 -> (i, j, k) = (1 + 5 % 4, 1 + (5 // 4) % 3, 1 + 5 // 12) = (2, 2, 1)
    # ordinal 4 -> (1, 2, 1): -2i clamps to i = 0; ordinal 7 -> (4, 2, 1): +2i clamps to i = 5
    for r, (i, j, k) in ((5, (2, 2, 1)), (4, (1, 2, 1)), (7, (4, 2, 1))):
        pts = stencil(i, j, k)
        c = 1  # u
        d = np.array([[qn[p, c] - qm1[p, c], qm1[p, c] - qm2[p, c]] for p in pts]).ravel()
        rms = np.sqrt(np.mean(d ** 2))
        assert np.allclose(x[r, 18 * c:18 * c + 18], d / rms, rtol=1e-6, atol=0), r
        # +j neighbour's d2 sits at column 18 c + 2 * 4 + 1
        assert np.isclose(x[r, 18 * c + 9], (qm1[pts[4], c] - qm2[pts[4], c]) / rms, rtol=1e-6)
        assert np.array_equal(x[r, 90:], qn[node(i, j, k)].astype(np.float32))
    # an unchanged history gives zero features, not a division by zero
    z = mo.features(qn, qn, qn, dims)
    assert np.array_equal(z[:, :90], np.zeros((24, 90), dtype=np.float32))


def test_zero_network_is_linear_extrapolation():
    dims = (5, 5, 5)
    n1 = 125
    rng = np.random.default_rng(9)
    q, qn, qm1, qm2 = (rng.normal(size=(n1, 5)) for _ in range(4))
    nh = 4
    w = np.zeros(mo.param_count(nh))
    stats = np.concatenate([np.zeros(95), np.ones(95), np.ones(10)])
    lin = mo.warm_start(q, qn, qm1, qm2, dims)
    net = mo.warm_start(q, qn, qm1, qm2, dims, w, nh, stats)
    g = mo.interior_index(dims)
    assert np.allclose(net[g], lin[g], rtol=0, atol=1e-12)
    assert np.array_equal(lin[g, 4], np.maximum(1e-7, 2 * qn[g, 4] - qm1[g, 4]))
    b = np.setdiff1d(np.arange(n1), g)
    assert np.array_equal(lin[b], q[b])
    # b = -1 with a = 2 is quadratic extrapolation 3 q^n - 3 q^(n-1) + q^(n-2)
    W3 = np.zeros((nh + 1, 10))
    W3[-1, :5] = 1.0
    W3[-1, 5:] = -1.0
    w2 = np.concatenate([np.zeros(((mo.N_IN + 1) * nh) + (nh + 1) * nh), W3.ravel()])
    quad = mo.warm_start(q, qn, qm1, qm2, dims, w2, nh, stats)
    want = 3 * qn[g] - 3 * qm1[g] + qm2[g]
    assert np.allclose(quad[g, :4], want[:, :4], rtol=0, atol=1e-12)
