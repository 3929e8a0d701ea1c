"""Overset grids: the donor search and cross-zone donor tables.

The reference reads its donor tables (global_osg.F90:18-28: for every
fringe node, the host cell's lower corner and zone and the trilinear weights)
from a file written by an external tool (init_osg_interface,
src/init.F90:1979-2000).  This module builds them itself:

  find_donors  for physical points, the host cell of a curvilinear block and
               the point's local (trilinear) coordinates in that cell, by
               Newton inversion of the cell's trilinear map, walking to the
               neighbouring cell while the solution leaves [0, 1]^3
  ZoneTable    a cross-zone table in the reference's order: hosts and nodes
               (n, 4) as (i, j, k, zone), 1-based; coef (n, 3)
  hill_blocks  a multi-block test geometry (blocks overlapping in x over a
               wavy bottom wall) with its fringe layers and donor table
"""
import math

import numpy as np
from scipy.spatial import cKDTree

from .cases import Case, fv1, tanh_stretch
from .params import make_params, BC_INTERFACE, BC_WALL, BC_PERIODIC


class ZoneTable:
    """Cross-zone donor table: hosts/nodes (n, 4) = (i, j, k, zone), 1-based;
    coef (n, 3) trilinear weights; apply_mom as for desalm_group_set_overset."""

    def __init__(self, hosts, coef, nodes, apply_mom=True):
        self.hosts = np.ascontiguousarray(hosts, dtype=np.int32).reshape(-1, 4)
        self.coef = np.ascontiguousarray(coef, dtype=np.float64).reshape(-1, 3)
        self.nodes = np.ascontiguousarray(nodes, dtype=np.int32).reshape(-1, 4)
        self.apply_mom = bool(apply_mom)

    def __len__(self):
        return len(self.coef)


def _corners(X, i, j, k):
    """the 8 corner positions (n, 8, 3) of cells with lower corners (i, j, k),
    0-based, in the order 000 100 010 001 101 011 110 111 of mg_bintp_mom"""
    off = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (0, 1, 1), (1, 1, 0), (1, 1, 1)]
    return np.stack([X[:, k + c, j + b, i + a].T for a, b, c in off], axis=1)


def _trilinear(cn, s):
    """positions and Jacobians of the trilinear maps of cells cn (n, 8, 3) at
    local coordinates s (n, 3)"""
    x, y, z = s[:, 0:1], s[:, 1:2], s[:, 2:3]
    w = [(1 - x) * (1 - y) * (1 - z), x * (1 - y) * (1 - z), (1 - x) * y * (1 - z), (1 - x) * (1 - y) * z,
         x * (1 - y) * z, (1 - x) * y * z, x * y * (1 - z), x * y * z]
    dx = [-(1 - y) * (1 - z), (1 - y) * (1 - z), -y * (1 - z), -(1 - y) * z, (1 - y) * z, -y * z, y * (1 - z), y * z]
    dy = [-(1 - x) * (1 - z), -x * (1 - z), (1 - x) * (1 - z), -(1 - x) * z, -x * z, (1 - x) * z, x * (1 - z), x * z]
    dz = [-(1 - x) * (1 - y), -x * (1 - y), -(1 - x) * y, (1 - x) * (1 - y), x * (1 - y), (1 - x) * y, -x * y, x * y]
    pos = sum(w[c] * cn[:, c, :] for c in range(8))
    jac = np.stack([sum(d[c] * cn[:, c, :] for c in range(8)) for d in (dx, dy, dz)], axis=2)
    return pos, jac


def find_donors(xyz, dims, pts, newton=20, walks=8, tol=1e-12):
    """Host cells of a structured block for physical points.

    xyz: (x, y, z) flat arrays of the block (i fastest); dims: (ni, nj, nk);
    pts: (n, 3).  Returns (hosts (n, 3) 1-based lower corners, coef (n, 3),
    found (n,) bool).  A point counts as found when its local coordinates lie
    in [0, 1]^3 (to 1e-9) and the mapped position matches it to tol times the
    block's extent."""
    ni, nj, nk = dims
    X = np.stack([np.asarray(a).reshape(nk, nj, ni) for a in xyz])
    pts = np.asarray(pts, dtype=np.float64)
    n = len(pts)
    tree = cKDTree(np.stack([np.asarray(a) for a in xyz], axis=1))
    _, near = tree.query(pts)
    i = np.clip(near % ni, 0, ni - 2)
    j = np.clip((near // ni) % nj, 0, nj - 2)
    k = np.clip(near // (ni * nj), 0, nk - 2)
    s = np.full((n, 3), 0.5)
    scale = max(float(np.ptp(X[d])) for d in range(3))
    found = np.zeros(n, dtype=bool)
    lo, hi = np.zeros(3, dtype=int), np.array([ni - 2, nj - 2, nk - 2])
    for _ in range(walks):
        todo = ~found
        if not todo.any():
            break
        cn = _corners(X, i[todo], j[todo], k[todo])
        st = np.full((todo.sum(), 3), 0.5)
        for _ in range(newton):
            pos, jac = _trilinear(cn, st)
            st = st - np.linalg.solve(jac, (pos - pts[todo])[:, :, None])[:, :, 0]
        pos, _ = _trilinear(cn, st)
        err = np.abs(pos - pts[todo]).max(axis=1)
        inside = np.all((st >= -1e-9) & (st <= 1 + 1e-9), axis=1)
        ok = inside & (err <= tol * scale)
        idx = np.nonzero(todo)[0]
        s[idx] = st
        found[idx[ok]] = True
        # walk towards the point: one cell per direction that left [0, 1]
        cur = np.stack([i[idx], j[idx], k[idx]], axis=1)
        step = np.where(st < 0, -1, np.where(st > 1, 1, 0))
        new = np.clip(cur + step, lo, hi)
        mv = ~ok
        i[idx[mv]], j[idx[mv]], k[idx[mv]] = new[mv, 0], new[mv, 1], new[mv, 2]
    s = np.clip(s, 0.0, 1.0)
    return np.stack([i + 1, j + 1, k + 1], axis=1), s, found


def hill_blocks(dims_list, x_splits, overlap=2, fringe=2, lz=1.0, amp=0.08, seed=11, des=True, ng=2, itm=3,
                iters=(1, 2, 3), ren=800.0, delti=0.02, apply_mom=True, **kw):
    """Blocks side by side in x over a wavy bottom wall y_w = amp sin(2 pi x)
    (a periodic-hill-like channel on [0, 1] x [y_w, 1] x [0, lz]).

    dims_list: (ni, nj, nk) per block; x_splits: the nominal interfaces between
    neighbouring blocks.  Each block reaches `overlap` of its own cells past an
    interface, and its last `fringe` i-planes on that side are fringe nodes
    filled from the neighbour.  Faces: bottom wall, top wall, periodic in k,
    interfaces btype 0 (left alone by bcond_fm), outer i faces walls.
    Returns (cases, ZoneTable)."""
    nb = len(dims_list)
    edges = [0.0] + list(x_splits) + [1.0]
    rng = np.random.default_rng(seed)
    blocks = []
    for b, (ni, nj, nk) in enumerate(dims_list):
        x0, x1 = edges[b], edges[b + 1]
        hx = (x1 - x0) / (ni - 1 - (overlap if b > 0 else 0) - (overlap if b < nb - 1 else 0))
        xa = x0 - (overlap * hx if b > 0 else 0.0)
        xs = xa + hx * np.arange(ni)
        eta = tanh_stretch(nj, 1.2)
        zs = np.linspace(0.0, lz, nk)
        Z, E, Xg = np.meshgrid(zs, eta, xs, indexing="ij")
        yw = amp * np.sin(2 * math.pi * Xg)
        Y = yw + (1.0 - yw) * E
        blocks.append((Xg.ravel().copy(), Y.ravel().copy(), Z.ravel().copy()))
    cases = []
    for b, (ni, nj, nk) in enumerate(dims_list):
        x, y, z = blocks[b]
        n1 = ni * nj * nk
        q = np.zeros((n1, 5))
        a = rng.uniform(-1, 1, 6)
        q[:, 0] = 0.1 * np.sin(3.0 * x + a[0]) * np.cos(2.0 * y + a[1])
        q[:, 1] = 4.0 * y * (1.0 - y) + 0.05 * np.sin(5.0 * z + a[2])
        q[:, 2] = 0.05 * np.sin(4.0 * x + a[3]) * np.sin(math.pi * y)
        q[:, 3] = 0.05 * np.cos(3.0 * x + a[4]) * np.sin(math.pi * y)
        q[:, 4] = (3.0 + np.sin(2.0 * x + a[5])) / ren
        qn = q.copy()
        qnm1 = q.copy()
        qn[:, :4] += 0.01 * rng.standard_normal((n1, 4))
        qnm1[:, :4] += 0.02 * rng.standard_normal((n1, 4))
        xnut = q[:, 4] * fv1(q[:, 4], ren)
        wd = np.minimum(y - amp * np.sin(2 * math.pi * x), 1.0 - y) + 1e-3
        bt = (BC_WALL if b == 0 else BC_INTERFACE, BC_WALL if b == nb - 1 else BC_INTERFACE,
              BC_WALL, BC_WALL, BC_PERIODIC, BC_PERIODIC)
        p = make_params(ni, nj, nk, ng=ng, itm=itm, iters=iters, ren=ren, delti=delti, unsteady=True, des=des,
                        btype=bt, fix=(3, 2, 4), **kw)
        cases.append(Case("hill_block", p, x, y, z, wd, q, qn, qnm1, xnut))
    # fringe nodes: the last `fringe` interior-side i-planes towards each
    # neighbour (interior j, k), filled from the neighbouring block
    hosts, coefs, nodes = [], [], []
    for b, (ni, nj, nk) in enumerate(dims_list):
        x, y, z = blocks[b]
        for side, nbr in ((0, b - 1), (1, b + 1)):
            if nbr < 0 or nbr >= nb:
                continue
            planes = range(1, fringe + 1) if side == 0 else range(ni - fringe + 1, ni + 1)
            for ip in planes:
                jj, kk = np.meshgrid(np.arange(2, nj), np.arange(2, nk), indexing="ij")
                jj, kk = jj.ravel(), kk.ravel()
                l = (ip - 1) + ni * ((jj - 1) + nj * (kk - 1))
                pts = np.stack([x[l], y[l], z[l]], axis=1)
                h, c, ok = find_donors(blocks[nbr], dims_list[nbr], pts)
                if not ok.all():
                    raise ValueError(f"block {b + 1}: {np.count_nonzero(~ok)} fringe points have no donor")
                hosts.append(np.column_stack([h, np.full(len(h), nbr + 1)]))
                coefs.append(c)
                nodes.append(np.column_stack([np.full(len(l), ip), jj, kk, np.full(len(l), b + 1)]))
    table = ZoneTable(np.concatenate(hosts), np.concatenate(coefs), np.concatenate(nodes), apply_mom)
    return cases, table
