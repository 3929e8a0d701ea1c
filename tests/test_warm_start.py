"""Learned warm start (desalm_ws_*, desalm_step) on the GPU against the CPU
restatement oracle/mlp_oracle.py, and what the trained network gains over
plain extrapolation in physical time."""
import numpy as np
import pytest

import mlp_oracle as mo
from desalm import cases
from desalm.lib import Solver

pytestmark = pytest.mark.gpu


def small_solver(nh=24, **kw):
    c = cases.wavy(13, 11, 9, ng=2, itm=2, iters=(1, 2), **kw)
    s = Solver(c.params).load_case(c)
    s.ws_configure(hidden=nh, samples_per_record=64, capacity=256, seed=3)
    return c, s


def test_mlp_forward_matches_oracle(gpu):
    nh = 24
    _, s = small_solver(nh)
    rng = np.random.default_rng(1)
    w = rng.normal(0.0, 0.2, mo.param_count(nh)).astype(np.float32)
    s.ws_params(w)
    assert np.array_equal(s.ws_params(), w)
    x = rng.normal(size=(3000, mo.N_IN)).astype(np.float32)  # not a multiple of the 64-row tile
    y = s.mlp_forward(x)
    yo = mo.forward(w.astype(np.float64), x.astype(np.float64), nh)
    # FP32 layers of at most 76 terms: a few ulps of the largest activation sum
    assert np.abs(y - yo).max() <= 2e-5 * max(1.0, np.abs(yo).max())
    s.close()


@pytest.mark.parametrize("m", [1, 700, 2500])
def test_mlp_gradient_matches_oracle(gpu, m):
    nh = 32
    _, s = small_solver(nh)
    rng = np.random.default_rng(m)
    w = rng.normal(0.0, 0.2, mo.param_count(nh)).astype(np.float32)
    s.ws_params(w)
    x = rng.normal(size=(m, mo.N_IN)).astype(np.float32)
    t = rng.normal(size=(m, 3 * mo.N_COMP)).astype(np.float32)
    loss, g = s.mlp_loss_grad(x, t)
    lo, go = mo.loss_grad(w.astype(np.float64), x.astype(np.float64), t.astype(np.float64), nh)
    assert abs(loss - lo) <= 1e-5 * abs(lo)
    # FP32 sums over up to 2500 rows (partials of 512 rows, then in order)
    assert np.abs(g - go).max() <= 1e-4 * np.abs(go).max()
    s.close()


def test_features_and_targets_match_oracle(gpu):
    c, s = small_solver()
    q, qn, qnm1, _ = s.get_state()
    rng = np.random.default_rng(11)
    q2 = qnm1 + 0.01 * rng.standard_normal(qnm1.shape)
    s.ws_history(q2)
    assert np.array_equal(s.ws_history(), q2)
    dims = c.params.level_dims()[0]
    m = (dims[0] - 2) * (dims[1] - 2) * (dims[2] - 2)
    x, t = s.ws_features(m)
    xo = mo.features(qn, qnm1, q2, dims)
    to = mo.targets(q, qn, qnm1, q2, dims).astype(np.float32)
    # the root mean square may be contracted into FMAs on the device: a few ulps
    assert np.allclose(x, xo, rtol=1e-6, atol=1e-7)
    assert np.array_equal(x[:, 90:], xo[:, 90:])
    assert np.array_equal(t, to)
    s.close()


def test_extrapolated_start_is_exact(gpu):
    c, s = small_solver()
    q, qn, qnm1, xnut = s.get_state()
    s.ws_mode(Solver.WS_EXTRAP)
    s.set_convergence(0, 0, -8.0, -30.0)  # the start only: no pseudo-iteration
    s.mg_driver()
    q1, qn1, qnm1_1, _ = s.get_state()
    # the shift qnm1 <- qn, qn <- q, then 2 q^n - q^(n-1) on the interior
    assert np.array_equal(qn1, q)
    assert np.array_equal(qnm1_1, qn)
    want = mo.warm_start(q, q, qn, qnm1, c.params.level_dims()[0])
    assert np.array_equal(q1, want)
    assert np.array_equal(s.ws_history(), qnm1)
    s.close()


def test_learned_start_matches_oracle(gpu):
    nh = 16
    c, s = small_solver(nh)
    q, qn, qnm1, _ = s.get_state()
    rng = np.random.default_rng(21)
    w = rng.normal(0.0, 0.3, mo.param_count(nh)).astype(np.float32)
    stats = np.concatenate([rng.normal(0.0, 0.3, 95), rng.uniform(0.5, 2.0, 95),
                            rng.uniform(0.01, 0.1, 5), np.zeros(5)]).astype(np.float32)
    s.ws_params(w)
    s.ws_stats(stats)
    s.ws_mode(Solver.WS_LEARNED)
    s.set_convergence(0, 0, -8.0, -30.0)
    s.mg_driver()
    q1 = s.get_state()[0]
    dims = c.params.level_dims()[0]
    want = mo.warm_start(q, q, qn, qnm1, dims, w.astype(np.float64), nh, stats.astype(np.float64))
    ext = mo.warm_start(q, q, qn, qnm1, dims)
    g = mo.interior_index(dims)
    corr_gpu = q1[g] - ext[g]
    corr_ref = want[g] - ext[g]
    # the coefficients are FP32 (features and network), applied in FP64
    scale = np.abs(corr_ref).max(axis=0)
    assert (scale > 0).all()
    assert (np.abs(corr_gpu - corr_ref).max(axis=0) <= 2e-5 * scale + 1e-14).all()
    s.close()


def test_step_advances_blade_angle(gpu):
    # main.F90:42-49: wrap past 2 pi first, then advance by TSR delti length_scale / tur_rad
    c = cases.alm_box(25, 17, 17, ng=2, itm=1, nelems=4, nangles=37, radius=0.5, hub=0.05)
    s = Solver(c.params).load_case(c)
    t = c.alm.turbines[0]
    d = t.tsr * c.params.delti * c.alm.length_scale / t.tur_rad
    s.turbine_angle(0, 1.0)
    s.step()
    assert s.turbine_angle(0) == 1.0 + d
    s.turbine_angle(0, 2 * np.pi + 0.25)
    s.step()
    assert s.turbine_angle(0) == (2 * np.pi + 0.25 - 2 * np.pi) + d
    s.close()


def test_trained_network_beats_extrapolation(gpu):
    """Train on 20 well converged plain steps of an actuator-line box, then
    start 6 unseen steps from linear extrapolation and from the learned
    coefficients: the learned starts must be closer to the converged
    solutions (the CPU study of this setup gives about 0.8 of the
    extrapolation error)."""
    c = cases.alm_box(41, 25, 25, ng=2, itm=150, nelems=6, nangles=73, iters=(1, 2))
    s = Solver(c.params).load_case(c)
    s.ws_configure(hidden=32, samples_per_record=8192, capacity=1 << 18, seed=5)
    dims = c.params.level_dims()[0]
    g = mo.interior_index(dims)
    for _ in range(3):          # history for the features
        s.step()
    for _ in range(20):
        s.step()
        s.ws_record()
    loss = s.ws_train(800, batch=4096, lr=2e-3)
    assert loss[-50:].mean() < 0.8 * loss[:10].mean()
    err = {Solver.WS_EXTRAP: [], Solver.WS_LEARNED: []}
    for _ in range(6):
        snap = s.snapshot()
        s.ws_mode(Solver.WS_NONE)
        s.set_convergence(150, 150, -99.0, -99.0)
        s.step()
        converged = s.get_state()[0]
        after = s.snapshot()
        for mode in err:
            s.restore(snap)
            s.ws_mode(mode)
            s.set_convergence(0, 0, -8.0, -30.0)
            s.step()
            start = s.get_state()[0]
            err[mode].append(np.abs(start[g, 1:4] - converged[g, 1:4]).mean())
        s.restore(after)
    ratio = np.array(err[Solver.WS_LEARNED]) / np.array(err[Solver.WS_EXTRAP])
    print("learned / extrapolation start error per unseen step:", np.round(ratio, 3))
    assert ratio.mean() < 0.95
    s.close()


def test_rfg_inflow_matches_oracle(gpu):
    """desalm_step applies the random-flow-generator inflow on the plane
    i = 1 at time0 + steps * delti before its V-cycles (main.F90:58)."""
    import rfg_oracle as ro
    from desalm import rfg
    c = cases.alm_box(25, 17, 17, ng=2, itm=1, rotor=False)
    s = Solver(c.params).load_case(c)
    dims = c.params.level_dims()[0]
    md = rfg.modes(64, 9)
    tb = rfg.inflow_tables(dims[1], dims[2], intensity=0.05, length=0.1)
    s.set_rfg(md, tb, time0=0.35)
    s.set_convergence(0, 0, -8.0, -30.0)  # the forcing only: no pseudo-iteration
    for step in range(2):
        q0 = s.get_state()[0]
        s.step()
        q1 = s.get_state()[0]
        want = ro.inflow(q0, (c.x, c.y, c.z), dims, 0.35 + step * c.params.delti, md, tb)
        g = np.arange(dims[1] * dims[2]) * dims[0]
        assert np.allclose(q1[g], want[g], rtol=1e-12, atol=1e-13)
        other = np.setdiff1d(np.arange(len(q0)), g)
        assert np.array_equal(q1[other], q0[other])
    s.close()
