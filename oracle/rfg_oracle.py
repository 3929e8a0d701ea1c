"""CPU restatement of the random-flow-generator inflow (test infrastructure
only; only tests/, __graft_entry__.smoke() and bench.py's CPU baseline may
use it).  Follows rfg.F90:329-374 (rfg_vel: scaled time and position,
isotropic mode sum, anti-diagonalisation) and rfg_inlet.F90:107-173 (the
interior nodes of the plane i = 1 get um + v), in float64.  Pinned by the
mode sum's statistics (mean 0, covariance equal to the Reynolds-stress
tensor over many samples), by the solenoidality of every mode (k . p =
k . q = 0) and against a per-node loop (tests/test_rfg_oracle.py)."""
import numpy as np


def rfg_vel(t, x, ts, ls, m, c, k, omega, p, q):
    """velocity fluctuation at points x (P, 3) with per-point ts, ls (P,),
    m (P, 9) column-major, c (P, 3); modes k (N, 3), omega (N,), p, q (N, 3)"""
    n = len(omega)
    t1 = t / ts
    x1 = x / ls[:, None]
    ups = ls / ts
    s = np.zeros((x.shape[0], 3))
    for i in range(n):
        th = np.zeros(x.shape[0])
        for d in range(3):
            th = th + k[i, d] * ups / c[:, d] * x1[:, d]
        th = th + omega[i] * t1
        cs, sn = np.cos(th), np.sin(th)
        for d in range(3):
            s[:, d] = s[:, d] + p[i, d] * cs + q[i, d] * sn
    cst = np.sqrt(2.0 / n)
    v1 = c * (cst * s)
    out = np.zeros_like(v1)
    for r in range(3):
        out[:, r] = m[:, r] * v1[:, 0] + m[:, r + 3] * v1[:, 1] + m[:, r + 6] * v1[:, 2]
    return out


def inflow(q, xyz, dims, t, modes, tables):
    """q (n1, 5) with the inflow plane's interior set as rfg_inlet does"""
    ni, nj, nk = dims
    k, omega, p, qq = modes
    um, m, c, ts, ls = tables
    out = q.copy()
    jj, kk = np.meshgrid(np.arange(1, nj - 1), np.arange(1, nk - 1), indexing="xy")
    jk = (jj + nj * kk).ravel()
    g = ni * jk
    x = np.stack([xyz[0][g], xyz[1][g], xyz[2][g]], axis=1)
    v = rfg_vel(t, x, ts[jk], ls[jk], m.reshape(-1, 9)[jk], c.reshape(-1, 3)[jk], k, omega, p, qq)
    out[g, 1:4] = um.reshape(-1, 3)[jk] + v
    return out
