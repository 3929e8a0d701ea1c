"""Closed-form and library checks of the oracle, independent of the
reference Fortran: each targets a term whose coefficients or indices a
restatement could get wrong."""
import numpy as np
import pytest
import scipy.linalg.lapack as lapack

from desalm import cases
from desalm.params import make_params
import pyoracle as po


@pytest.mark.parametrize("n", [1, 2, 3, 7, 50])
def test_gtsv_equals_lapack_dgtsv(n):
    rng = np.random.default_rng(n)
    for trial in range(20):
        dl = rng.uniform(-1, 1, max(n - 1, 1))[: n - 1]
        d = rng.uniform(-1, 1, n)
        du = rng.uniform(-1, 1, max(n - 1, 1))[: n - 1]
        b = rng.standard_normal((n, 3))
        x, info = po.gtsv(dl, d, du, b)
        if n >= 2:
            ref = lapack.dgtsv(dl, d, du, b)
            assert info == ref[-1]
            if info == 0:
                assert np.array_equal(x, ref[3])  # same algorithm, same rounding
        else:
            assert info == 0 and np.allclose(x[0], b[0] / d[0])


def test_gtsv_zero_pivot_info():
    d = np.array([0.0, 1.0, 1.0])
    dl = np.array([0.0, 0.5])
    du = np.array([1.0, 0.5])
    _, info = po.gtsv(dl, d, du, np.ones(3))
    assert info == 1


def _affine_case(ni=9, nj=8, nk=7, A=None):
    A = np.array([[1.3, 0.2, 0.1], [0.15, 0.9, -0.2], [0.05, 0.1, 1.1]]) if A is None else A
    I, J, K = np.meshgrid(np.arange(ni), np.arange(nj), np.arange(nk), indexing="ij")
    xi = np.stack([I.transpose(2, 1, 0).ravel(), J.transpose(2, 1, 0).ravel(), K.transpose(2, 1, 0).ravel()])
    X = A @ xi
    p = make_params(ni, nj, nk, ng=1, btype=(0, 0, 0, 0, 0, 0))
    n1 = ni * nj * nk
    z = np.zeros(n1)
    q = np.zeros((n1, 5))
    return cases.Case("affine", p, X[0].copy(), X[1].copy(), X[2].copy(), z, q, q.copy(), q.copy(), z.copy()), A


def test_metrics_affine_grid_exact():
    c, A = _affine_case()
    o = po.Oracle(c.params).load_case(c)
    Ainv = np.linalg.inv(A)
    # csi = grad xi = first row of A^-1, etc.; J = 1 / det(A)
    assert np.allclose(o.get("csi"), Ainv[0], atol=1e-14)
    assert np.allclose(o.get("eta"), Ainv[1], atol=1e-14)
    assert np.allclose(o.get("zet"), Ainv[2], atol=1e-14)
    assert np.allclose(o.get("aj"), 1.0 / np.linalg.det(A), rtol=1e-14)


def test_free_stream_preserved():
    c, _ = _affine_case(11, 10, 9)
    c.q[:, 0] = 0.7
    c.q[:, 1:4] = (0.9, -0.4, 0.3)
    c.qn[:] = c.q
    c.qnm1[:] = c.q
    c.params.unsteady = 1
    o = po.Oracle(c.params).load_case(c)
    rh = o.residual_only(1)
    assert np.max(np.abs(rh)) < 1e-12
    o.solver_daf(1, 0, 0)
    q = o.get("q")
    # velocities unchanged; pressure shifted by the pinned value (bcond_fm.F90:310-322)
    assert np.max(np.abs(q[:, 1:4] - c.q[:, 1:4])) < 1e-12
    assert np.max(np.abs(q[:, 0])) < 1e-12


def _uniform_case(n=11, h=0.1):
    s = np.arange(n) * h
    x, y, z = cases.mesh(s, s, s)
    p = make_params(n, n, n, ng=1, btype=(0, 0, 0, 0, 0, 0), unsteady=False, pdiss=0.0, ren=50.0)
    n1 = n ** 3
    q = np.zeros((n1, 5))
    zz = np.zeros(n1)
    return cases.Case("uniform", p, x, y, z, zz, q, q.copy(), q.copy(), zz.copy()), x, y, z


def test_quick_exact_for_quadratics():
    """QUICK (coef 1/8; 3,3,-7,1 weights) is exact for quadratic fields, so on a
    uniform grid with U along x the convective part of rh(2) equals
    (U/J) * d(u)/d(csi) at every node that is not first or last interior."""
    n, h = 11, 0.1
    c, x, y, z = _uniform_case(n, h)
    U = 1.7
    c.q[:, 1] = U            # transport velocity u = const ...
    c.q[:, 3] = 0.3 * x ** 2  # ... of the quadratic w(x): convected by u
    c.params.ren = 1e12      # viscous term negligible
    o = po.Oracle(c.params).load_case(c)
    rh = o.residual_only(1)
    aj = 1.0 / h ** 3
    # U/J = (u / h) h^3 = u h^2 and dw/dcsi = h dw/dx, so the term is u h^3 dw/dx
    expect = U * h ** 3 * 0.6 * x
    idx = np.arange(n ** 3)
    i, j, k = idx % n, (idx // n) % n, idx // n ** 2
    inner = (i >= 2) & (i <= n - 3) & (j >= 1) & (j <= n - 2) & (k >= 1) & (k <= n - 2)
    assert np.max(np.abs(rh[inner, 3] - expect[inner])) < 1e-12 * np.max(np.abs(expect))
    assert aj > 0


def test_viscous_closed_form_quadratic():
    """Stress-form viscous term on a uniform grid (spacing h, J = h^-3): the
    x-momentum term is visc(1) = nu h^3 [lap(u) + d/dx(div u)], exact for
    quadratic velocity fields (rhs_viscous.F90:158-163 carries the transpose
    term a1 uu1 + a2 vv1 + a3 ww1)."""
    n, h = 9, 0.2
    c, x, y, z = _uniform_case(n, h)
    nu = 1.0 / c.params.ren
    c.q[:, 1] = 0.5 * x ** 2 + 0.25 * y ** 2 - z ** 2  # no convection: v = w = 0, u does not move along x...
    c.q[:, 1] -= 0.0
    o = po.Oracle(c.params).load_case(c)
    # convection of u by itself is non-zero; remove it by comparing against a
    # run with nu -> 0
    rh = o.residual_only(1)
    c2, *_ = _uniform_case(n, h)
    c2.params.ren = 1e300
    c2.q[:] = c.q
    o2 = po.Oracle(c2.params).load_case(c2)
    rh0 = o2.residual_only(1)
    visc = -(rh[:, 1] - rh0[:, 1])
    lap = (0.5 * 2 + 0.25 * 2 - 2.0)  # d2u/dx2 + d2u/dy2 + d2u/dz2
    grad_div = 1.0                     # div u = du/dx = x, d/dx of it = 1
    expect = nu * h ** 3 * (lap + grad_div)
    idx = np.arange(n ** 3)
    i, j, k = idx % n, (idx // n) % n, idx // n ** 2
    inner = (i >= 1) & (i <= n - 2) & (j >= 1) & (j <= n - 2) & (k >= 1) & (k <= n - 2)
    assert np.allclose(visc[inner], expect, rtol=1e-10, atol=1e-15)


def test_pressure_dissipation_zero_for_cubic():
    n, h = 11, 0.1
    c, x, y, z = _uniform_case(n, h)
    c.params.pdiss_coef = 0.3
    c.q[:, 0] = x ** 3 - 2 * y ** 3 + 0.5 * z ** 3
    o = po.Oracle(c.params).load_case(c)
    rh_on = o.residual_only(1)
    c.params.pdiss_coef = 0.0
    o2 = po.Oracle(c.params).load_case(c)
    rh_off = o2.residual_only(1)
    idx = np.arange(n ** 3)
    i, j, k = idx % n, (idx // n) % n, idx // n ** 2
    inner = (i >= 2) & (i <= n - 3) & (j >= 2) & (j <= n - 3) & (k >= 2) & (k <= n - 3)
    # 4th difference of a cubic vanishes away from the one-sided boundary copies
    assert np.max(np.abs(rh_on[inner, 0] - rh_off[inner, 0])) < 1e-12


def test_restriction_preserves_constants_and_linears():
    c = cases.wavy(17, 13, 11, ng=2)
    o = po.Oracle(c.params).load_case(c)
    nall = o.nall
    n1 = c.n1
    idx = np.arange(n1)
    i, j, k = idx % 17, (idx // 17) % 13, idx // (17 * 13)
    rh = np.zeros((nall, 4))
    rh[:n1, 0] = 2.5
    rh[:n1, 1] = 0.3 * i - 0.2 * j + 0.1 * k
    o.set("rh", rh)
    o.mg_nlevel(2)
    pk = o.get("pk")[n1:]
    ci = np.arange(9 * 7 * 6)
    ii, jj, kk = ci % 9, (ci // 9) % 7, ci // 63
    inner = (ii > 0) & (ii < 8) & (jj > 0) & (jj < 6) & (kk > 0) & (kk < 5)
    assert np.allclose(pk[inner, 0], 2.5, rtol=1e-15)
    lin = 0.3 * (2 * ii) - 0.2 * (2 * jj) + 0.1 * (2 * kk)
    assert np.allclose(pk[inner, 1], lin[inner], atol=1e-13)
    assert np.all(pk[~inner] == 0.0)


def test_prolongation_exact_for_linear_corrections():
    c = cases.wavy(17, 13, 11, ng=2, ep=((1.0, 2.5, 2.5), (0.0, 0.0, 0.0)))
    o = po.Oracle(c.params).load_case(c)
    n1 = c.n1
    q = o.get("q")
    qold = o.get("qold_mg")
    ci = np.arange(9 * 7 * 6)
    ii, jj, kk = ci % 9, (ci // 9) % 7, ci // 63
    # coarse correction linear in the fine index space
    corr = 0.01 * (2 * ii) + 0.02 * (2 * jj) - 0.03 * (2 * kk)
    qold[n1:, :4] = 0.0
    q[n1:, :4] = corr[:, None]
    o.set("q", q)
    o.set("qold_mg", qold)
    before = o.get("q")[:n1, :4].copy()
    o.mg_prolong(1)
    dq = o.get("q")[:n1, :4] - before
    idx = np.arange(n1)
    i, j, k = idx % 17, (idx // 17) % 13, idx // (17 * 13)
    inner = (i > 0) & (i < 16) & (j > 0) & (j < 12) & (k > 0) & (k < 10)
    lin = 0.01 * i + 0.02 * j - 0.03 * k
    assert np.allclose(dq[inner, 2], lin[inner], atol=1e-14)
    assert np.all(dq[~inner] == 0.0)


def test_no_row_interchanges_in_workloads():
    """The DAF, SA-ADI and smoothing systems of the workloads never make
    SGTSV interchange rows, so a pivot-free elimination follows its order."""
    po.gtsv_pivots(reset=True)
    c = cases.channel(33, 33, 33, ng=3, itm=3, iters=(1, 2, 3))
    po.Oracle(c.params).load_case(c).mg_driver()
    c = cases.cavity(17, ng=2, itm=5)
    po.Oracle(c.params).load_case(c).mg_driver()
    assert po.gtsv_pivots(reset=True) == 0


def test_bintp_exact_for_trilinear_fields():
    """mg_bintp_mom/tur reproduce a field that is trilinear in the node
    indices: the value at a fringe node is the field at host + coef."""
    c = cases.wavy(13, 11, 9, ng=2, des=True, overset=50, seed=17)
    ni, nj, nk = 13, 11, 9
    k, j, i = np.meshgrid(np.arange(1, nk + 1), np.arange(1, nj + 1), np.arange(1, ni + 1), indexing="ij")
    i, j, k = i.ravel().astype(float), j.ravel().astype(float), k.ravel().astype(float)

    def f(a, b, d, s):
        return s * (0.3 + 0.2 * a - 0.1 * b + 0.05 * d + 0.01 * a * b - 0.02 * b * d + 0.015 * a * d
                    + 0.001 * a * b * d)

    q = c.q.copy()
    for m in range(5):
        q[:, m] = f(i, j, k, 1.0 + m)
    q[:, 4] += 10.0  # nu~ > 0 keeps the floor of mg_bintp_tur inactive
    ov = c.overset
    pts = ov.hosts + ov.coef
    expect = np.stack([f(pts[:, 0], pts[:, 1], pts[:, 2], 1.0 + m) for m in range(5)], axis=1)
    expect[:, 4] += 10.0
    fringe = (ov.nodes[:, 0] - 1) + ni * ((ov.nodes[:, 1] - 1) + nj * (ov.nodes[:, 2] - 1))
    for which, comps in (("mom", range(4)), ("tur", [4])):
        o = po.Oracle(c.params).load_case(c)
        o.set_state(q=q)
        o.bintp(which)
        out = o.get_state()[0]
        for m in comps:
            assert np.allclose(out[fringe, m], expect[:, m], rtol=1e-13, atol=1e-13), (which, m)
        untouched = np.setdiff1d(np.arange(c.n1), fringe)
        assert np.array_equal(out[untouched], q[untouched])
