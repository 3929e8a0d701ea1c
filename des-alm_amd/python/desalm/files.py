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


# ---- text outputs of a physical step (src/main.F90:71-75) ----

def fortran_g(x, w, d):
    """x written with the Fortran edit descriptor Gw.d (nearest rounding):
    F editing with d significant digits followed by four blanks when the
    rounded magnitude lies in [0.1, 10**d), E editing otherwise; zero as
    F(w-4).(d-1) followed by four blanks."""
    x = float(x)
    if x != x:
        return "NaN".rjust(w)
    if x in (float("inf"), float("-inf")):
        return ("-Infinity" if x < 0 else "Infinity").rjust(w)
    neg = x < 0 or (x == 0.0 and str(x).startswith("-"))

    def fit(body, width):
        s = ("-" if neg else "") + body
        if len(s) > width and body.startswith("0."):
            s = ("-" if neg else "") + body[1:]
        return "*" * width if len(s) > width else s.rjust(width)

    if x == 0.0:
        return fit("%.*f" % (d - 1, 0.0), w - 4) + "    "
    mant, ex = ("%.*e" % (d - 1, abs(x))).split("e")
    e10 = int(ex)          # exponent of the value rounded to d digits
    s = e10 + 1            # 10**(s-1) <= rounded value < 10**s
    if 0 <= s <= d:
        body = "%.*f" % (d - s, abs(x))
        return fit(body if d > s else body + ".", w - 4) + "    "
    digits = mant.replace(".", "")
    e = e10 + 1
    exp = ("E%+03d" % e) if abs(e) <= 99 else ("%+04d" % e)
    return fit("0." + digits + exp, w)


def conver_real_line(ntime, residual):
    """'conver_real' line, format (i5.5,1x,5(g13.6,1x)) (src/main.F90:199)."""
    return "%05d " % ntime + "".join(fortran_g(v, 13, 6) + " " for v in residual)


def history_line(ntime, values):
    """'history' line of one monitor point: q(1:5) and xnut, format
    (i5.5,1x,6(g12.6,1x)) (src/main.F90:552)."""
    return "%05d " % ntime + "".join(fortran_g(v, 12, 6) + " " for v in values)


def ctcp_line(ntime, ct, cp):
    """'resultsCTCPnn.dat' line, format (i5.5,2(1x,G20.7)) (src/main.F90:103)."""
    return "%05d" % ntime + " " + fortran_g(ct, 20, 7) + " " + fortran_g(cp, 20, 7)


def append_step_outputs(directory, ntime, solver, monitor_points=(), n_turbines=0):
    """Append the reference's per-step text outputs after a physical step
    (src/main.F90:71-75): the real residual to 'conver_real'; q(1:5) and
    xnut at every monitor point (1-based fine-level i, j, k) to 'history';
    and per rotor CT and CP to 'resultsCTCPnn.dat'.  The reference's
    actuator-line path sets C_T = C_P = 1 (src/rhs_lines.F90:458-459, the
    thrust and power expressions are commented out there), so those are
    the values written."""
    import os
    with open(os.path.join(directory, "conver_real"), "a") as f:
        f.write(conver_real_line(ntime, solver.real_residual()) + "\n")
    if len(monitor_points):
        vals = solver.probe(monitor_points)
        with open(os.path.join(directory, "history"), "a") as f:
            for v in vals:
                f.write(history_line(ntime, v) + "\n")
    for t in range(n_turbines):
        with open(os.path.join(directory, "resultsCTCP%02d.dat" % (t + 1)), "a") as f:
            f.write(ctcp_line(ntime, 1.0, 1.0) + "\n")
