"""Random-flow-generator inflow tables (desalm_set_rfg of desalm.h).

The method is that of rfg.F90 (Smirnov, Shi & Celik, J. Fluids Eng. 123
(2001) 359-371): N random Fourier modes with wavenumbers k_n and frequencies
omega_n drawn from normal distributions (standard deviations 0.5 and 1),
amplitudes p_n = zeta_n x k_n, q_n = xi_n x k_n (zeta_n, xi_n standard
normal), rescaled per node by the eigen-decomposition of the Reynolds-stress
tensor (rfg.F90:329-374).  The reference draws the modes with its own
uniform generator and a polar Box-Muller transform (rfg.F90:112-200); here
they come from numpy's PCG64 generator, seeded.  `inflow_tables` builds the
per-node tables of the inflow plane from a mean profile, an intensity and
an integral length.
"""
import numpy as np


def modes(n, seed):
    """(k (n, 3), omega (n,), p (n, 3), q (n, 3)) of n random modes"""
    rng = np.random.Generator(np.random.PCG64(seed))
    zeta = rng.normal(0.0, 1.0, (n, 3))
    xi = rng.normal(0.0, 1.0, (n, 3))
    k = rng.normal(0.0, 0.5, (n, 3))
    omega = rng.normal(0.0, 1.0, n)
    return k, omega, np.cross(zeta, k), np.cross(xi, k)


def eig_tables(r):
    """per node: eigenvectors (column-major 3 x 3, each column's entry of
    largest magnitude positive) and square roots of the eigenvalues, in
    ascending order, of the symmetric tensors r (..., 3, 3)"""
    w, v = np.linalg.eigh(r)
    idx = np.argmax(np.abs(v), axis=-2)                       # (..., 3) row of the largest entry per column
    sgn = np.sign(np.take_along_axis(v, idx[..., None, :], axis=-2))
    v = v * np.where(sgn == 0, 1.0, sgn)
    c = np.sqrt(np.maximum(w, 0.0))
    m = np.swapaxes(v, -1, -2).reshape(v.shape[:-2] + (9,))   # column-major: m[r + 3 col] = v[r, col]
    return m, c


def inflow_tables(jmg, kmg, um=(1.0, 0.0, 0.0), intensity=0.05, length=0.1, u_ref=1.0, anisotropy=(1.0, 1.0, 1.0)):
    """(um (3 jmg kmg), m (9 ...), c (3 ...), ts, ls) for a uniform mean
    flow with turbulence of the given intensity and integral length on every
    node of the inflow plane (index j + jmg k); ts = length / u_ref"""
    nplane = jmg * kmg
    var = (intensity * u_ref) ** 2 * np.asarray(anisotropy, dtype=float)
    r = np.zeros((nplane, 3, 3))
    for d in range(3):
        r[:, d, d] = var[d]
    m, c = eig_tables(r)
    umean = np.tile(np.asarray(um, dtype=float), (nplane, 1))
    ts = np.full(nplane, length / u_ref)
    ls = np.full(nplane, float(length))
    return umean, m, c, ts, ls
