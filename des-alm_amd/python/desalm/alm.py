"""Actuator-line model tables (desalm_turbine / desalm_polar of desalm.h).

The reference builds these tables in init_lines (src/init.F90:1156-1710)
from its grid file, a blade description (blade.txt) and airfoil polar files
read during the run (src/rhs_lines.F90:232-300).  `rotor` builds the same
kind of tables for any block: a rotor of nblades blades of nelems elements
in the plane i = i_rot, elements binned by blade angle in nangles bins,
host cells found with the overset donor search, disk data (radius, area,
volume, chord, twist, polar blend, thickness) per rotor-plane node, the
spreading nodes around the disk and the hub nodes.
"""
import ctypes
import math

import numpy as np

from .overset import find_donors

i32p = ctypes.POINTER(ctypes.c_int32)
dp = ctypes.POINTER(ctypes.c_double)


class TurbineC(ctypes.Structure):
    _fields_ = [("nblades", ctypes.c_int32), ("nelems", ctypes.c_int32), ("nangles", ctypes.c_int32),
                ("itur1", ctypes.c_int32), ("itur2", ctypes.c_int32),
                ("tur_rad", ctypes.c_double), ("hub_rad", ctypes.c_double), ("tsr", ctypes.c_double),
                ("angle", ctypes.c_double), ("dtheta_interp", ctypes.c_double),
                ("node_indexes", i32p), ("node_indexes_pos", dp), ("disk_dat", dp),
                ("polar_before", i32p), ("polar_after", i32p),
                ("n_disk", ctypes.c_int32), ("disk_nodes", i32p),
                ("n_ker", ctypes.c_int32), ("ker_nodes", i32p), ("ker_pos", dp), ("ker_dat", dp)]


class PolarC(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("angle", dp), ("cl", dp), ("cd", dp)]


class Turbine:
    """One rotor's tables (numpy, layouts as documented in desalm.h)."""

    def __init__(self, **kw):
        self.__dict__.update(kw)


class AlmSetup:
    """Turbines, polars and the scales of module global_turb."""

    def __init__(self, turbines, polars, length_scale, vel_infty):
        self.turbines = list(turbines)
        self.polars = list(polars)  # (angle, cl, cd) arrays
        self.length_scale = float(length_scale)
        self.vel_infty = float(vel_infty)

    def as_ctypes(self):
        """(TurbineC array, PolarC array, keep-alive list)"""
        keep = []

        def ip(a):
            a = np.ascontiguousarray(a, dtype=np.int32)
            keep.append(a)
            return a.ctypes.data_as(i32p)

        def fp(a):
            a = np.ascontiguousarray(a, dtype=np.float64)
            keep.append(a)
            return a.ctypes.data_as(dp)
        ts = (TurbineC * len(self.turbines))()
        for t, T in enumerate(self.turbines):
            ts[t] = TurbineC(T.nblades, T.nelems, T.nangles, T.itur1, T.itur2, T.tur_rad, T.hub_rad, T.tsr,
                             T.angle, T.dtheta_interp, ip(T.node_indexes), fp(T.node_indexes_pos),
                             fp(T.disk_dat), ip(T.polar_before), ip(T.polar_after), len(T.disk_nodes),
                             ip(T.disk_nodes), len(T.ker_nodes), ip(T.ker_nodes), fp(T.ker_pos), fp(T.ker_dat))
        ps = (PolarC * max(len(self.polars), 1))()
        for k, (a, cl, cd) in enumerate(self.polars):
            ps[k] = PolarC(len(a), fp(a), fp(cl), fp(cd))
        return ts, ps, keep


def polars(n=3):
    """n smooth synthetic polars over -180..180 degrees (2-degree steps): a
    thin-airfoil slope up to stall, a flat-plate branch beyond it."""
    a = np.arange(-180.0, 180.0 + 1e-9, 2.0)
    out = []
    for k in range(n):
        stall = 12.0 + 2.0 * k
        r = np.radians(a)
        cl_lin = 2 * math.pi * np.sin(r) * (1.0 - 0.1 * k)
        cl_fp = 1.1 * np.sin(2 * r)
        w = 1.0 / (1.0 + np.exp((np.abs(a) - stall) / 2.0))
        cl = w * cl_lin + (1 - w) * cl_fp
        cd = 0.008 + 0.002 * k + (1 - w) * 1.2 * np.sin(r) ** 2 + 0.01 * (a / 20.0) ** 2 * w
        out.append((a, cl, cd))
    return out


def rotor(case, i_rot, center, radius, hub, nblades=3, nelems=8, nangles=73, tsr=6.0, angle=0.3,
          length_scale=1.0, ker_reach=3, n_polar=3):
    """Tables of one rotor on the fine level of `case` (a desalm.cases.Case):
    rotor plane i = i_rot (1-based), axis along the i direction through
    center = (yc, zc), blade radius `radius`, hub radius `hub`.  vel_infty is
    set to ren 1e-6 / length_scale, the velocity scale of rhs_lines, so that
    the non-dimensional inflow is one."""
    p = case.params
    ni, nj, nk = p.img, p.jmg, p.kmg
    X = case.x.reshape(nk, nj, ni)
    Y = case.y.reshape(nk, nj, ni)
    Z = case.z.reshape(nk, nj, ni)
    yc, zc = center
    xr = float(np.mean(X[:, :, i_rot - 1]))
    dtheta = 2 * math.pi / (nangles - 1)
    # elements: radial stations and angle bins (angle bin b covers b * dtheta)
    rel = hub + (np.arange(nelems) + 0.5) * (radius - hub) / nelems
    pts, order = [], []
    for b in range(nangles):
        th = b * dtheta
        for e in range(nelems):
            pts.append((xr, yc + rel[e] * math.sin(th), zc + rel[e] * math.cos(th)))
            order.append((e, b))
    pts = np.array(pts)
    h, s, ok = find_donors((case.x, case.y, case.z), (ni, nj, nk), pts)
    if not ok.all():
        raise ValueError("rotor elements outside the block")
    node_indexes = np.zeros((3, nangles, nelems), dtype=np.int32)
    node_pos = np.zeros((27, nangles, nelems))
    corner = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    for idx, (e, b) in enumerate(order):
        i, j, k = (int(v) for v in h[idx])
        node_indexes[:, b, e] = (i, j, k)
        node_pos[0:3, b, e] = pts[idx]
        for c, (a1, b1, c1) in enumerate(corner):
            ii, jj, kk = i - 1 + a1, j - 1 + b1, k - 1 + c1
            node_pos[3 + 3 * c: 6 + 3 * c, b, e] = (X[kk, jj, ii], Y[kk, jj, ii], Z[kk, jj, ii])
    # disk data per (j, k) of the rotor plane (disk_dat(j,k,1:8))
    yy, zz = Y[:, :, i_rot - 1].T, Z[:, :, i_rot - 1].T  # (nj, nk)
    rad = np.hypot(yy - yc, zz - zc)
    dy = np.gradient(yy, axis=0)
    dz = np.gradient(zz, axis=1)
    dx = np.gradient(X[:, :, :], axis=2)[:, :, i_rot - 1].T
    area = np.abs(dy * dz)
    vol = np.abs(area * dx)
    frac = np.clip((rad - hub) / (radius - hub), 0.0, 1.0)
    chord = (0.12 - 0.06 * frac) * radius
    twist = 20.0 - 16.0 * frac
    pos = frac * (n_polar - 1)
    before = np.minimum(np.floor(pos).astype(np.int32), n_polar - 1)
    after = np.minimum(before + 1, n_polar - 1)
    inter = pos - before
    relthick = 0.25 - 0.1 * frac
    disk = np.stack([rad / length_scale, area / length_scale ** 2, np.arctan2(zz - zc, yy - yc),
                     vol / length_scale ** 3, chord, twist, inter, relthick], axis=2)  # (nj, nk, 8)
    # spreading nodes: interior nodes within ker_reach cells of the rotor disk
    ker, kpos, kdat = [], [], []
    for k in range(2, nk):
        for j in range(2, nj):
            for i in range(max(2, i_rot - ker_reach), min(ni - 1, i_rot + ker_reach) + 1):
                x, y, z = X[k - 1, j - 1, i - 1], Y[k - 1, j - 1, i - 1], Z[k - 1, j - 1, i - 1]
                r = math.hypot(y - yc, z - zc)
                if r <= radius * 1.2:
                    ker.append((i, j, k))
                    kpos.append((x, y, z))
                    kdat.append((r / length_scale, area[j - 1, k - 1] / length_scale ** 2,
                                 vol[j - 1, k - 1] / length_scale ** 3))
    disk_nodes = [(i_rot, j, k) for k in range(1, nk + 1) for j in range(1, nj + 1) if rad[j - 1, k - 1] <= radius]
    T = Turbine(nblades=nblades, nelems=nelems, nangles=nangles, itur1=i_rot - 1, itur2=i_rot + 1,
                tur_rad=radius, hub_rad=hub, tsr=tsr, angle=angle, dtheta_interp=dtheta,
                node_indexes=node_indexes.ravel(), node_indexes_pos=node_pos.ravel(),
                disk_dat=np.transpose(disk, (2, 1, 0)).ravel(),
                polar_before=before.T.ravel(), polar_after=after.T.ravel(),
                disk_nodes=np.array(disk_nodes, dtype=np.int32).reshape(-1, 3),
                ker_nodes=np.array(ker, dtype=np.int32).reshape(-1, 3), ker_pos=np.array(kpos).reshape(-1, 3),
                ker_dat=np.array(kdat).reshape(-1, 3))
    vel_infty = p.ren * 1e-6 / length_scale
    return AlmSetup([T], polars(n_polar), length_scale, vel_infty)
