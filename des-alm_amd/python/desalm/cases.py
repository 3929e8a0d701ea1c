"""Synthetic workloads: grids, wall distance and initial fields.

Arrays are numpy, flat over the fine level with the reference's node order
(i fastest, then j, then k; src/init.F90:880-980).  Vector fields are
(n1, m) C-ordered, i.e. the reference's q(1:m, l) layout.

Workloads (constants not stated by the workload definition take the shipped
deck's values, ind3dmg.dat:8-16):
  cavity   33^3 lid-driven cavity, Re 1000, steady, laminar, 2 levels
  channel  129^3 SA channel, Re_tau 395, log-law start + 5% seeded noise,
           3 levels, BDF2 dual time
  alm_box  193x129x129 SA-DES box, uniform inflow U = 1 past one actuator
           line rotor (3 blades, R = 0.5, TSR 6), dt = 0.01
"""
import math

import numpy as np

from .params import make_params, BC_WALL, BC_INTERFACE, BC_PERIODIC, BC_PERIODIC_WRAP, BC_SYMMETRY, BC_FREESTREAM, BC_INFLOW


class Overset:
    """Overset donor table of the fine level (src/global_osg.F90:18-28).

    Donor d interpolates trilinearly from the cell whose lower corner is
    hosts[d] (1-based i, j, k) at weights coef[d] into the fringe node
    nodes[d] (1-based i, j, k).  nu~ is interpolated after every des_eddy
    (mg_driver.F90:182); q(1:4) after every fine smoother pass when
    apply_mom is set, which the reference does when nzone > 1
    (mg_driver.F90:241-244)."""

    def __init__(self, hosts, coef, nodes, apply_mom=True):
        self.hosts = np.ascontiguousarray(hosts, dtype=np.int32).reshape(-1, 3)
        self.coef = np.ascontiguousarray(coef, dtype=np.float64).reshape(-1, 3)
        self.nodes = np.ascontiguousarray(nodes, dtype=np.int32).reshape(-1, 3)
        self.apply_mom = bool(apply_mom)


def overset_table(ni, nj, nk, nd, seed=5, reach=2, apply_mom=True):
    """nd distinct interior fringe nodes, each fed by a host cell within
    `reach` cells of it, with uniform random trilinear weights."""
    rng = np.random.default_rng(seed)
    n_int = (ni - 2) * (nj - 2) * (nk - 2)
    pick = rng.choice(n_int, size=nd, replace=False)
    i = 2 + pick % (ni - 2)
    j = 2 + (pick // (ni - 2)) % (nj - 2)
    k = 2 + pick // ((ni - 2) * (nj - 2))
    nodes = np.stack([i, j, k], axis=1)
    off = rng.integers(-reach, reach + 1, size=(nd, 3))
    lim = np.array([ni - 1, nj - 1, nk - 1])
    hosts = np.clip(nodes + off, 1, lim)
    coef = rng.uniform(0.0, 1.0, size=(nd, 3))
    return Overset(hosts, coef, nodes, apply_mom)


class Case:
    """A complete input set for one run of the inner iteration."""

    def __init__(self, name, params, x, y, z, wd, q, qn, qnm1, xnut, overset=None):
        self.name = name
        self.params = params
        self.x, self.y, self.z = x, y, z
        self.wd = wd
        self.q, self.qn, self.qnm1, self.xnut = q, qn, qnm1, xnut
        self.overset = overset
        self.alm = None  # desalm.alm.AlmSetup: actuator-line rotors

    @property
    def n1(self):
        return self.params.img * self.params.jmg * self.params.kmg

    def state(self):
        return self.q.copy(), self.qn.copy(), self.qnm1.copy(), self.xnut.copy()


def tanh_stretch(n, delta):
    """x = 1/2 [1 + tanh(delta (2 s - 1)) / tanh(delta)], s uniform on [0,1]."""
    s = np.linspace(0.0, 1.0, n)
    return 0.5 * (1.0 + np.tanh(delta * (2.0 * s - 1.0)) / math.tanh(delta))


def mesh(xi, yj, zk):
    """Tensor-product coordinates flattened in i-fastest order."""
    Z, Y, X = np.meshgrid(zk, yj, xi, indexing="ij")
    return X.ravel().copy(), Y.ravel().copy(), Z.ravel().copy()


def fv1(nut, ren):
    chi = nut * ren
    chic = chi ** 3
    return chic / (chic + 357.911)


def cavity(n=33, ng=2, itm=200, stretch=1.1, ren=1000.0, **kw):
    """Plain FP64 laminar lid-driven cavity (workload 1)."""
    s = tanh_stretch(n, stretch)
    x, y, z = mesh(s, s, s)
    bt = [BC_WALL] * 6
    bt[3] = BC_INTERFACE  # lid j = jmax: nodes left untouched, u = 1 preserved
    p = make_params(n, n, n, ng=ng, itm=itm, ren=ren, beta=1.0, unsteady=False, des=False,
                    btype=bt, fix=(2, 2, 2), **kw)
    n1 = n * n * n
    q = np.zeros((n1, 5))
    j = (np.arange(n1) // n) % n
    q[j == n - 1, 1] = 1.0
    zero1 = np.zeros(n1)
    return Case("cavity", p, x, y, z, zero1.copy(), q, q.copy(), q.copy(), zero1.copy())


def channel_y(nj, dy1_wall):
    """Wall-normal tanh-stretched coordinate on [-1, 1] whose first spacing
    at each wall is dy1_wall (solved for the stretching factor)."""
    s = np.linspace(0.0, 1.0, nj)

    def first(g):
        yy = np.tanh(g * (2.0 * s - 1.0)) / math.tanh(g)
        return yy[1] - yy[0]

    lo, hi = 1e-3, 10.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if first(mid) > dy1_wall:
            lo = mid
        else:
            hi = mid
    g = 0.5 * (lo + hi)
    return np.tanh(g * (2.0 * s - 1.0)) / math.tanh(g)


def loglaw(yplus):
    return np.where(yplus < 11.0, yplus, np.log(np.maximum(yplus, 1e-30)) / 0.41 + 5.2)


def periodic_axis(n, length):
    """Uniform nodes of a periodic direction: nodes 1 .. n-2 (0-based) cover
    one period [0, length) and the end planes are the images of the opposite
    interior planes, x(0) = x(n-2) - length, x(n-1) = x(1) + length."""
    h = length / (n - 2)
    return (np.arange(n) - 1.0) * h


def channel(ni=129, nj=129, nk=129, ng=3, itm=30, ren=395.0, delti=0.005, seed=20260101,
            noise=0.05, des=True, periodic="wrap", **kw):
    """Turbulent channel, SA (workload 2): Re_tau = ren with u_tau = 1 and
    half-height 1, walls at j = 1 and j = jmax, domain 2 pi x 2 x pi.

    periodic = "wrap": streamwise and spanwise periodic (btype 7, whose end
    planes are images of the opposite interior planes) and the mean pressure
    gradient as a body force f_x = u_tau^2 / h = 1.  periodic = "copy": the
    reference's btype 6 (the end planes copy each other and freeze) and no
    body force, on the same grid spacing."""
    if periodic not in ("wrap", "copy"):
        raise ValueError("periodic must be 'wrap' or 'copy'")
    xi = periodic_axis(ni, 2.0 * math.pi)
    yj = channel_y(nj, 2.0 / ren if nj < 65 else 1.0 / ren)
    zk = periodic_axis(nk, math.pi)
    x, y, z = mesh(xi, yj, zk)
    n1 = ni * nj * nk
    wd = 1.0 - np.abs(y)
    yplus = wd * ren
    u = loglaw(yplus)
    uc = float(loglaw(np.array([ren]))[0])
    rng = np.random.default_rng(seed)
    q = np.zeros((n1, 5))
    q[:, 1] = u + noise * uc * rng.uniform(-1.0, 1.0, n1)
    q[:, 2] = noise * uc * rng.uniform(-1.0, 1.0, n1)
    q[:, 3] = noise * uc * rng.uniform(-1.0, 1.0, n1)
    wall = (wd == 0.0)
    q[wall, 1:4] = 0.0
    # nu~ from a mixing-length estimate, clipped to the SA floor of 1e-7
    nut = np.maximum(0.41 * wd * np.minimum(yplus, 200.0) / ren * (1.0 - np.exp(-yplus / 26.0)), 1.0e-7)
    q[:, 4] = nut
    xnut = nut * fv1(nut, ren)
    per = BC_PERIODIC_WRAP if periodic == "wrap" else BC_PERIODIC
    bt = [per, per, BC_WALL, BC_WALL, per, per]
    fb = (1.0, 0.0, 0.0) if periodic == "wrap" else (0.0, 0.0, 0.0)
    if periodic == "wrap":
        # the end planes start as the images of the opposite interior planes
        for a in (q,):
            a3 = a.reshape(nk, nj, ni, 5)
            a3[:, :, 0] = a3[:, :, ni - 2]
            a3[:, :, ni - 1] = a3[:, :, 1]
            a3[0] = a3[nk - 2]
            a3[nk - 1] = a3[1]
        xnut = q[:, 4] * fv1(q[:, 4], ren)
    p = make_params(ni, nj, nk, ng=ng, itm=itm, ren=ren, beta=1.0, delti=delti, unsteady=True,
                    des=des, btype=bt, fix=(2, nj // 2, 2), body_force=fb, **kw)
    return Case("channel", p, x, y, z, wd, q, q.copy(), q.copy(), xnut)


def laminar_channel(ni=9, nj=17, nk=9, ng=2, itm=20, ren=10.0, force=1.0, **kw):
    """Steady laminar plane Poiseuille flow: walls at y = -1 and 1 (uniform
    spacing), periodic-wrap i and k, body force (force, 0, 0).  Its exact
    solution u = force ren (1 - y^2) / 2, v = w = p = 0 is also the exact
    discrete steady state (central differences are exact for a quadratic on
    a uniform grid), and the returned case starts from it."""
    xi = periodic_axis(ni, 1.0)
    yj = np.linspace(-1.0, 1.0, nj)
    zk = periodic_axis(nk, 0.5)
    x, y, z = mesh(xi, yj, zk)
    n1 = ni * nj * nk
    q = np.zeros((n1, 5))
    q[:, 1] = 0.5 * force * ren * (1.0 - y * y)
    q[np.abs(y) == 1.0, 1] = 0.0
    wd = 1.0 - np.abs(y)
    xnut = np.zeros(n1)
    bt = [BC_PERIODIC_WRAP, BC_PERIODIC_WRAP, BC_WALL, BC_WALL, BC_PERIODIC_WRAP, BC_PERIODIC_WRAP]
    p = make_params(ni, nj, nk, ng=ng, itm=itm, ren=ren, beta=1.0, unsteady=False, des=False, btype=bt,
                    fix=(2, nj // 2, 2), body_force=(force, 0.0, 0.0), **kw)
    return Case("laminar_channel", p, x, y, z, wd, q, q.copy(), q.copy(), xnut)


def wavy(ni, nj, nk, amp=0.04, ng=2, seed=7, des=False, unsteady=True, btype=None, ren=500.0,
         itm=3, delti=0.02, stretch=0.8, overset=0, apply_mom=True, **kw):
    """Curvilinear test block: a stretched box with smooth sinusoidal
    distortion of every coordinate (non-orthogonal metrics, all cross terms
    non-zero) and smooth random initial fields.  Used by the parity tests.
    overset > 0 adds a donor table of that many fringe nodes."""
    si = tanh_stretch(ni, stretch)
    sj = tanh_stretch(nj, stretch)
    sk = tanh_stretch(nk, stretch)
    X, Y, Z = mesh(si, sj, sk)
    two_pi = 2.0 * math.pi
    x = X + amp * np.sin(two_pi * Y) * np.sin(math.pi * Z)
    y = Y + amp * np.sin(two_pi * Z) * np.sin(math.pi * X)
    z = Z + amp * np.sin(two_pi * X) * np.sin(math.pi * Y)
    n1 = ni * nj * nk
    rng = np.random.default_rng(seed)

    def smooth(scale):
        a = rng.uniform(-1, 1, 6)
        return scale * (a[0] * np.sin(3.1 * X + a[1]) * np.cos(2.3 * Y + a[2]) +
                        a[3] * np.sin(1.7 * Z + a[4]) + 0.3 * a[5] * X * Y)

    q = np.zeros((n1, 5))
    q[:, 0] = smooth(0.3)
    q[:, 1] = 1.0 + smooth(0.5)
    q[:, 2] = smooth(0.4)
    q[:, 3] = smooth(0.4)
    wd = np.minimum(np.minimum(Y, 1.0 - Y), np.minimum(Z, 1.0 - Z)) + 1e-3
    q[:, 4] = (2.0 + smooth(1.0) + 1.5) / ren
    qn = q.copy()
    qnm1 = q.copy()
    qn[:, :4] += 0.01 * rng.standard_normal((n1, 4))
    qnm1[:, :4] += 0.02 * rng.standard_normal((n1, 4))
    qn[:, 4] *= 1.0 + 0.01 * rng.standard_normal(n1)
    qnm1[:, 4] *= 1.0 + 0.02 * rng.standard_normal(n1)
    xnut = q[:, 4] * fv1(q[:, 4], ren)
    if btype is None:
        btype = (BC_WALL, 3, BC_WALL, BC_INTERFACE, 2, 3)
    p = make_params(ni, nj, nk, ng=ng, itm=itm, ren=ren, delti=delti, unsteady=unsteady, des=des,
                    btype=btype, fix=(3, 2, 4), **kw)
    ov = overset_table(ni, nj, nk, overset, apply_mom=apply_mom) if overset else None
    return Case("wavy", p, x, y, z, wd, q, qn, qnm1, xnut, overset=ov)


def cluster(n, a, b, c, alpha):
    """n points on [a, b] clustered around c: x(s) = s - alpha/(2 pi)
    sin(2 pi (s - s_c)), monotonic for alpha < 1, spacing (1 - alpha) h at c
    and (1 + alpha) h far from it."""
    s = np.linspace(0.0, 1.0, n)
    sc = (c - a) / (b - a)
    x = s - alpha / (2.0 * math.pi) * np.sin(2.0 * math.pi * (s - sc))
    x = (x - x[0]) / (x[-1] - x[0])
    return a + (b - a) * x


def alm_box(ni=193, nj=129, nk=129, ng=3, itm=30, delti=0.01, ren=1.0e6, radius=0.5, tsr=6.0, hub=0.05,
            length=6.0, half=1.5, x_rot=1.5, nblades=3, nelems=16, nangles=None, angle=0.0, seed=20260404,
            noise=0.0, des=True, iters=(1, 2, 3), rotor=True, **kw):
    """Uniform inflow U = 1 past one actuator-line rotor (workload 4): an
    x = [0, length] by y, z = [-half, half] box clustered around the rotor
    plane x = x_rot and the rotor axis, inflow at i = 1 (held), zero-gradient
    outflow at i = imax, slip side faces.  SA-DES: no walls, so the SA
    length scale wd is the DES one, 0.65 times the largest local spacing.
    The rotor's tables are built on the fine level (desalm.alm.rotor); the
    Gaussian width is 2 dV^(1/3), i.e. two cells (rhs_lines.F90:434)."""
    xi = cluster(ni, 0.0, length, x_rot, 0.6)
    yj = cluster(nj, -half, half, 0.0, 0.6)
    zk = cluster(nk, -half, half, 0.0, 0.6)
    x, y, z = mesh(xi, yj, zk)
    n1 = ni * nj * nk
    dx = np.gradient(xi)
    dy = np.gradient(yj)
    dz = np.gradient(zk)
    dmax = np.maximum(np.maximum(dx[None, None, :], dy[None, :, None]), dz[:, None, None]).ravel()
    wd = 0.65 * dmax
    q = np.zeros((n1, 5))
    q[:, 1] = 1.0
    if noise > 0.0:
        rng = np.random.default_rng(seed)
        inner = np.ones(n1, dtype=bool)
        ii = np.arange(n1) % ni
        inner &= ii > 0
        q[:, 1] += noise * rng.uniform(-1.0, 1.0, n1) * inner
        q[:, 2] += noise * rng.uniform(-1.0, 1.0, n1) * inner
        q[:, 3] += noise * rng.uniform(-1.0, 1.0, n1) * inner
    q[:, 4] = 3.0 / ren
    xnut = q[:, 4] * fv1(q[:, 4], ren)
    bt = (BC_INFLOW, BC_FREESTREAM, BC_SYMMETRY, BC_SYMMETRY, BC_SYMMETRY, BC_SYMMETRY)
    p = make_params(ni, nj, nk, ng=ng, itm=itm, ren=ren, beta=1.0, delti=delti, unsteady=True, des=des,
                    btype=bt, fix=(ni - 1, nj // 2, nk // 2), iters=iters, **kw)
    c = Case("alm_box", p, x, y, z, wd, q, q.copy(), q.copy(), xnut)
    if rotor:
        from . import alm
        i_rot = int(np.argmin(np.abs(xi - x_rot))) + 1
        if nangles is None:
            nangles = 361
        c.alm = alm.rotor(c, i_rot=i_rot, center=(0.0, 0.0), radius=radius, hub=hub, nblades=nblades,
                          nelems=nelems, nangles=nangles, tsr=tsr, angle=angle, length_scale=1.0)
    return c
