"""The periodic-wrap boundary (btype 7) and the body force: the project's
channel forcing, which the reference does not have (its btype 6 copies the
end planes onto each other, so they freeze: src/bcond_fm.F90:197-202).

Pinned by what the mathematics fixes, on the oracle:
- plane Poiseuille flow, u = f Re (1 - y^2) / 2 between walls at y = -1, 1,
  periodic in x and z and driven by the body force f, is the exact discrete
  steady state on a uniform grid, so pseudo-iterations from it leave it in
  place to rounding: a wrong sign or scale of the force (-f / J in rh), a
  missing 1 / Re or a wrong periodic image moves it;
- a perturbation of that flow decays back towards it;
- after every pseudo-iteration both end planes of a wrapped direction hold
  the images of the opposite interior planes, for p, u, v, w and nu~, while
  btype 6 keeps its frozen plane copies.
Then the HIP path against the oracle at 1e-10 on the SA channel with the
wrap and the body force (tests/test_gpu_configs.py covers the copy form and
the 129^3 size).
"""
import numpy as np
import pytest

from desalm import cases
import pyoracle as po


def _run(c):
    o = po.Oracle(c.params).load_case(c)
    itc, h = o.mg_driver()
    n1 = c.q.shape[0]
    q, _, _, xnut = o.get_state()
    return itc, h, q[:n1], xnut[:n1]


def test_poiseuille_is_a_fixed_point():
    c = cases.laminar_channel(itm=20)
    itc, h, q, _ = _run(c)
    assert itc == 20
    umax = 0.5 * c.params.body_force[0] * c.params.ren
    assert np.max(np.abs(q[:, 1] - c.q[:, 1])) < 1e-12 * umax
    assert np.max(np.abs(q[:, 0:4] - c.q[:, 0:4])) < 1e-12 * umax
    # the update norms stay at the rounding level
    assert np.all(h[:, [0, 2, 3]] < -14.0)


def test_poiseuille_without_the_force_is_not_steady():
    c = cases.laminar_channel(itm=5)
    c.params.body_force[0] = 0.0
    _, _, q, _ = _run(c)
    # the flow decelerates everywhere inside
    inner = np.abs(c.y) < 0.99
    assert np.all(q[inner, 1] < c.q[inner, 1])


def test_perturbed_poiseuille_returns():
    c = cases.laminar_channel(itm=400)
    n1 = c.q.shape[0]
    rng = np.random.default_rng(3)
    inner = (np.abs(c.y) < 1.0)[:, None]
    pert = 0.05 * np.sin(np.pi * (c.y + 1.0) / 2.0)[:, None] * rng.uniform(0.5, 1.5, (n1, 3))
    exact = c.q.copy()
    q0 = exact.copy()
    q0[:, 1:4] = q0[:, 1:4] + pert * inner
    c.q, c.qn, c.qnm1 = q0, q0.copy(), q0.copy()
    _, _, q, _ = _run(c)
    e0 = np.max(np.abs(q0[:, 1:4] - exact[:, 1:4]))
    e1 = np.max(np.abs(q[:, 1:4] - exact[:, 1:4]))
    assert e1 < 0.35 * e0, (e0, e1)


def _planes(a, dims):
    ni, nj, nk = dims
    return a.reshape(nk, nj, ni, *a.shape[1:])


@pytest.mark.parametrize("periodic", ["wrap", "copy"])
def test_end_planes(periodic):
    c = cases.channel(17, 17, 17, ng=2, itm=3, iters=(1, 2), periodic=periodic)
    dims = (17, 17, 17)
    q0 = _planes(c.q.copy(), dims)
    _, _, q, xnut = _run(c)
    Q = _planes(q, dims)
    if periodic == "wrap":
        for m in range(5):
            assert np.array_equal(Q[:, :, 0, m], Q[:, :, 15, m]), m
            assert np.array_equal(Q[:, :, 16, m], Q[:, :, 1, m]), m
            assert np.array_equal(Q[0, :, :, m], Q[15, :, :, m]), m
            assert np.array_equal(Q[16, :, :, m], Q[1, :, :, m]), m
        # and they moved with the interior
        assert not np.array_equal(Q[:, :, 0, 1], q0[:, :, 0, 1])
    else:
        # the reference's copy: the velocities of the i planes are left as
        # they are (p moves with the pinned-pressure shift of every node),
        # the k planes copy each other
        assert np.array_equal(Q[1:-1, 1:-1, 0, 1:4], q0[1:-1, 1:-1, 0, 1:4])
        assert np.array_equal(Q[0, :, :, :4], Q[16, :, :, :4])


def test_wrap_needs_both_faces():
    # refused by the parameter check, before any device call
    from desalm.lib import DesalmError, Solver

    c = cases.channel(17, 17, 17, ng=2, itm=1)
    c.params.btype[1] = 1
    with pytest.raises(DesalmError, match="btype 7 on both"):
        Solver(c.params)


@pytest.mark.gpu
@pytest.mark.parametrize("n", [33])
def test_wrap_channel_gpu_matches_oracle(gpu, n):
    from desalm.lib import Solver

    c = cases.channel(n, n, n, ng=3, itm=3, iters=(1, 2, 3), periodic="wrap")
    o = po.Oracle(c.params).load_case(c)
    s = Solver(c.params).load_case(c)
    for _ in range(2):
        itc_o, ho = o.mg_driver()
        itc_g, hg = s.mg_driver()
        assert itc_o == itc_g
        assert np.max(np.abs(hg - ho)) < 1e-9
    qg, _, _, xg = s.get_state()
    qo, _, _, xo = o.get_state()
    for m in range(5):
        assert po.relerr(qg[:, m], qo[:, m]) < 1e-10, m
    assert po.relerr(xg, xo) < 1e-10
    assert s.pivot_events == 0
    s.close()


@pytest.mark.gpu
def test_poiseuille_fixed_point_gpu(gpu):
    from desalm.lib import Solver

    c = cases.laminar_channel(ni=17, nj=33, nk=17, itm=20)
    s = Solver(c.params).load_case(c)
    itc, h = s.mg_driver()
    q = s.get_state()[0][: c.q.shape[0]]
    umax = 0.5 * c.params.body_force[0] * c.params.ren
    assert np.max(np.abs(q[:, 0:4] - c.q[:, 0:4])) < 1e-12 * umax
    s.close()
