"""The reference's files through libdesalm's host-side readers and writers
(include/desalm.h, "The reference's files"): the grid file and the solution
file (Fortran unformatted sequential, REAL*4 or REAL*8), the restart output,
filestat and the input deck ind3dmg.dat.  No device is used.

Arrays follow the C ABI: x, y, z, wd and xnut of shape (n,), q and qn of
shape (n, 5), nodes i fastest, then j, then k, zones one after the other.
"""
import ctypes

import numpy as np

from .lib import DesalmError, lib
from .params import MAX_LEVELS, Params

MAX_TURBINES = 32


class DeckTurbine(ctypes.Structure):
    _fields_ = [
        ("itur1", ctypes.c_int32), ("itur2", ctypes.c_int32),
        ("jtur1", ctypes.c_int32), ("jtur2", ctypes.c_int32),
        ("ktur1", ctypes.c_int32), ("ktur2", ctypes.c_int32),
        ("tur_rad", ctypes.c_double), ("hub_rad", ctypes.c_double), ("tsr", ctypes.c_double),
        ("nblades", ctypes.c_int32), ("nelems", ctypes.c_int32),
    ]


class Deck(ctypes.Structure):
    """ctypes mirror of desalm_deck: the content of ind3dmg.dat
    (src/init.F90:176-228)."""
    _fields_ = [
        ("params", Params),
        ("ns", ctypes.c_int32), ("icycle", ctypes.c_int32), ("nt2", ctypes.c_int32),
        ("icn", ctypes.c_int32), ("icnw", ctypes.c_int32), ("irk", ctypes.c_int32),
        ("checkpoint", ctypes.c_int32), ("nzfix", ctypes.c_int32),
        ("itm", ctypes.c_int32 * MAX_LEVELS),
        ("alfa", ctypes.c_double * 4), ("eps", ctypes.c_double * 4),
        ("cdes", ctypes.c_double), ("ckin", ctypes.c_double), ("cein", ctypes.c_double),
        ("monitor_num_points", ctypes.c_int32),
        ("monitor_point_ijk", (ctypes.c_int32 * 4) * 10),
        ("length_scale", ctypes.c_double), ("vel_infty", ctypes.c_double),
        ("turbines", ctypes.c_int32), ("node_indexes_load", ctypes.c_int32),
        ("n_tur", ctypes.c_int32),
        ("turbine", DeckTurbine * MAX_TURBINES),
        ("nzblanking", ctypes.c_int32),
        ("blktype", ctypes.c_int32 * 6),
        ("blanking", (ctypes.c_int32 * 2) * 3),
        ("reserved", ctypes.c_int32 * 8),
    ]


_bound = False


def _lib():
    global _bound
    L = lib()
    if not _bound:
        dp = ctypes.POINTER(ctypes.c_double)
        i32 = ctypes.c_int32
        i32p = ctypes.POINTER(i32)
        cs = ctypes.c_char_p
        L.desalm_read_grid_file.argtypes = [cs, i32, i32p, i32, i32, dp, dp, dp, dp]
        L.desalm_read_solu_file.argtypes = [cs, i32, i32p, i32, i32, dp, dp, dp]
        L.desalm_write_solu_file.argtypes = [cs, ctypes.c_int64, i32, i32, dp, dp, dp]
        L.desalm_read_filestat.argtypes = [cs, i32, dp, i32p, dp]
        L.desalm_write_filestat.argtypes = [cs, ctypes.c_double, i32, i32, dp]
        L.desalm_read_deck.argtypes = [cs, i32, ctypes.POINTER(Deck)]
        for f in ("read_grid_file", "read_solu_file", "write_solu_file", "read_filestat", "write_filestat",
                  "read_deck"):
            getattr(L, "desalm_" + f).restype = ctypes.c_int
        _bound = True
    return L


def _chk(rc, what):
    if rc != 0:
        msg = lib().desalm_last_error(None)
        raise DesalmError(msg.decode() if msg else f"{what}: error {rc}")


def _dp(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _dims(dims):
    d = np.ascontiguousarray(np.asarray(dims, dtype=np.int32).reshape(-1, 3))
    return d, int(d.shape[0]), int(np.prod(d, axis=1).sum())


def read_grid(path, dims, turbulence=True, real_bytes=0):
    """x, y, z (and wd when turbulence, else None) of the grid file."""
    d, nz, n = _dims(dims)
    x, y, z = (np.empty(n) for _ in range(3))
    wd = np.empty(n) if turbulence else None
    _chk(_lib().desalm_read_grid_file(str(path).encode(), nz, d.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                      int(bool(turbulence)), real_bytes, _dp(x), _dp(y), _dp(z), _dp(wd)),
         "desalm_read_grid_file")
    return x, y, z, wd


def read_solu(path, dims, turbulence=True, real_bytes=0):
    """q (n, 5), qn (n, 5) and xnut (n,) (None without turbulence)."""
    d, nz, n = _dims(dims)
    q = np.empty((n, 5))
    qn = np.empty((n, 5))
    xnut = np.empty(n) if turbulence else None
    _chk(_lib().desalm_read_solu_file(str(path).encode(), nz, d.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                      int(bool(turbulence)), real_bytes, _dp(q), _dp(qn), _dp(xnut)),
         "desalm_read_solu_file")
    return q, qn, xnut


def write_solu(path, q, qn, xnut=None, real_bytes=4):
    """The restart output of one rank: q, xnut (when given), qn."""
    q = np.ascontiguousarray(q, dtype=np.float64)
    qn = np.ascontiguousarray(qn, dtype=np.float64)
    if q.ndim != 2 or q.shape[1] != 5 or qn.shape != q.shape:
        raise ValueError("q and qn must both have shape (n, 5)")
    if xnut is not None:
        xnut = np.ascontiguousarray(xnut, dtype=np.float64)
        if xnut.shape != (q.shape[0],):
            raise ValueError("xnut must have shape (n,)")
    _chk(_lib().desalm_write_solu_file(str(path).encode(), q.shape[0], int(xnut is not None), real_bytes,
                                       _dp(q), _dp(qn), _dp(xnut)), "desalm_write_solu_file")


def read_filestat(path, n_tur):
    """(time0, nt1, angles) of filestat."""
    t = ctypes.c_double()
    n = ctypes.c_int32()
    ang = np.zeros(max(n_tur, 1))
    _chk(_lib().desalm_read_filestat(str(path).encode(), n_tur, ctypes.byref(t), ctypes.byref(n), _dp(ang)),
         "desalm_read_filestat")
    return t.value, n.value, ang[:n_tur]


def write_filestat(path, time, ntime, angles=()):
    a = np.ascontiguousarray(np.asarray(angles, dtype=np.float64).reshape(-1))
    _chk(_lib().desalm_write_filestat(str(path).encode(), float(time), int(ntime), a.size,
                                      _dp(a) if a.size else None), "desalm_write_filestat")


def read_deck(path, ng):
    """Deck of ind3dmg.dat for a build with ng multigrid levels."""
    d = Deck()
    _chk(_lib().desalm_read_deck(str(path).encode(), ng, ctypes.byref(d)), "desalm_read_deck")
    return d
