"""Pins the CPU restatement of the random-flow-generator inflow
(oracle/rfg_oracle.py, after rfg.F90:329-374): every mode is solenoidal, the
generated fluctuations have zero mean and the prescribed Reynolds-stress
covariance over many samples, the eigen tables reproduce the tensor, and the
vectorised sum equals a per-point loop."""
import numpy as np

import rfg_oracle as ro
from desalm import rfg


def test_modes_are_solenoidal():
    k, omega, p, q = rfg.modes(500, 3)
    assert np.abs(np.einsum("nd,nd->n", k, p)).max() < 1e-12
    assert np.abs(np.einsum("nd,nd->n", k, q)).max() < 1e-12


def test_eig_tables_reproduce_the_tensor():
    rng = np.random.default_rng(4)
    a = rng.normal(size=(7, 3, 3))
    r = a @ np.swapaxes(a, 1, 2) + 0.1 * np.eye(3)
    m, c = rfg.eig_tables(r)
    M = m.reshape(7, 3, 3).swapaxes(1, 2)  # back from column-major
    rec = M @ (c[:, :, None] ** 2 * np.swapaxes(M, 1, 2))
    assert np.allclose(rec, r, rtol=1e-12, atol=1e-12)
    assert (np.diff(c, axis=1) >= 0).all()


def test_statistics_match_the_reynolds_stress():
    # many modes, many (t, x) samples: mean -> 0, covariance -> r
    k, omega, p, q = rfg.modes(400, 11)
    r = np.array([[0.04, 0.01, 0.0], [0.01, 0.02, 0.005], [0.0, 0.005, 0.03]])
    m, c = rfg.eig_tables(r[None])
    rng = np.random.default_rng(5)
    P = 4000
    x = rng.uniform(-20.0, 20.0, (P, 3))
    ts = np.ones(P)
    ls = np.ones(P)
    v = ro.rfg_vel(rng.uniform(0, 50.0), x, ts, ls, np.repeat(m, P, 0), np.repeat(c, P, 0), k, omega, p, q)
    cov = np.cov(v.T)
    assert np.abs(v.mean(axis=0)).max() < 0.02
    assert np.abs(cov - r).max() < 0.1 * np.abs(r).max()
    # a plausible slip (the eigenvalues used unsquared-rooted) fails the check
    v_bad = ro.rfg_vel(1.0, x, ts, ls, np.repeat(m, P, 0), np.repeat(c ** 2, P, 0), k, omega, p, q)
    assert np.abs(np.cov(v_bad.T) - r).max() > 0.5 * np.abs(r).max()


def test_vector_sum_matches_point_loop():
    k, omega, p, q = rfg.modes(9, 2)
    rng = np.random.default_rng(6)
    x = rng.normal(size=(4, 3))
    ts = rng.uniform(0.5, 2, 4)
    ls = rng.uniform(0.5, 2, 4)
    r = np.diag([0.01, 0.02, 0.03])
    m, c = rfg.eig_tables(np.repeat(r[None], 4, 0))
    v = ro.rfg_vel(0.7, x, ts, ls, m, c, k, omega, p, q)
    for i in range(4):
        M = m[i].reshape(3, 3).T
        s = np.zeros(3)
        for n in range(9):
            th = np.dot(k[n] * (ls[i] / ts[i]) / c[i], x[i] / ls[i]) + omega[n] * 0.7 / ts[i]
            s += p[n] * np.cos(th) + q[n] * np.sin(th)
        want = M @ (c[i] * (np.sqrt(2.0 / 9) * s))
        assert np.allclose(v[i], want, rtol=1e-12, atol=1e-14)
