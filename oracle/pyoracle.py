"""TEST INFRASTRUCTURE: Python bindings of the CPU oracle (liboracle.so) and a
runner for the reference Fortran harness (oracle/_ref/desalm_ref_harness).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this module, and only as the checker of the HIP path.
"""
import ctypes
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# DESALM_ORACLE_LIB: another build of the same sources (the sanitizer build
# of the Makefile's `sanitized` target)
LIB_PATH = os.environ.get("DESALM_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")
REF_EXE = os.path.join(HERE, "_ref", "desalm_ref_harness")
# the same harness with the reference's conservation and wf_mpi linked in
HOOKS_EXE = os.path.join(HERE, "_ref", "desalm_ref_harness_hooks")
HOOK_CONSERVATION = 1
HOOK_WALL_FUNCTIONS = 2
SHIM_MGD_EXE = os.path.join(HERE, "_ref", "desalm_shim_mgd")
SHIM_DAF_EXE = os.path.join(HERE, "_ref", "desalm_shim_daf")
# solver_daf drop-in inside the reference built with its case hooks
SHIM_DAF_HOOKS_EXE = os.path.join(HERE, "_ref", "desalm_shim_daf_hooks")

_lib = None


def build():
    """Compile liboracle.so (gcc) if it is missing or stale."""
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        dp = ctypes.POINTER(ctypes.c_double)
        L.or_create.restype = vp
        L.or_create.argtypes = [vp]
        L.or_destroy.argtypes = [vp]
        L.or_nall.restype = ctypes.c_long
        L.or_nall.argtypes = [vp]
        L.or_level_offset.restype = ctypes.c_long
        L.or_level_offset.argtypes = [vp, ctypes.c_int]
        for f in ("or_set_grid",):
            getattr(L, f).argtypes = [vp, dp, dp, dp]
        L.or_set_wd.argtypes = [vp, dp]
        L.or_set_state.argtypes = [vp, dp, dp, dp, dp]
        L.or_get_state.argtypes = [vp, dp, dp, dp, dp]
        L.or_get_array.argtypes = [vp, ctypes.c_char_p, dp]
        L.or_set_array.argtypes = [vp, ctypes.c_char_p, dp]
        L.or_mg_driver.argtypes = [vp, ctypes.POINTER(ctypes.c_int), dp]
        L.or_solver_daf.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.or_des_eddy.argtypes = [vp]
        L.or_residual.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.or_residual.restype = ctypes.c_int
        L.or_dtau.argtypes = [vp, ctypes.c_int, ctypes.c_void_p]
        L.or_dtau.restype = ctypes.c_int
        L.or_gtsv_pivots.argtypes = [ctypes.c_int]
        L.or_gtsv_pivots.restype = ctypes.c_long
        L.or_mg_nlevel.argtypes = [vp, ctypes.c_int]
        L.or_mg_prolong.argtypes = [vp, ctypes.c_int]
        L.or_mg_residual.argtypes = [vp, dp]
        L.or_gtsv.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp, dp, dp, ctypes.c_int]
        ip = ctypes.POINTER(ctypes.c_int)
        L.or_set_overset.argtypes = [vp, ctypes.c_long, ip, dp, ip, ctypes.c_int]
        L.or_set_hooks.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.or_wf_root.argtypes = [ctypes.c_double] * 3
        L.or_wf_root.restype = ctypes.c_double
        L.or_wf_newt.argtypes = [dp] + [ctypes.c_double] * 3
        L.or_group_mg_driver.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.c_long, ip, dp, ip, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_int), dp]
        L.or_group_bintp.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.c_long, ip, dp, ip, ctypes.c_int]
        L.or_set_turbines.argtypes = [vp, ctypes.c_int, vp, ctypes.c_int, vp, ctypes.c_double, ctypes.c_double]
        L.or_bintp.argtypes = [vp, ctypes.c_int]
        for f in ("or_set_grid", "or_set_wd", "or_set_state", "or_get_state", "or_get_array",
                  "or_set_array", "or_mg_driver", "or_solver_daf", "or_des_eddy", "or_mg_nlevel",
                  "or_mg_prolong", "or_mg_residual", "or_gtsv", "or_set_overset", "or_bintp", "or_group_mg_driver", "or_group_bintp", "or_set_turbines"):
            getattr(L, f).restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


NCOMP = {"q": 5, "qn": 5, "qnm1": 5, "qold_mg": 5, "rh": 4, "pk": 4, "csi": 3, "eta": 3,
         "zet": 3, "aj": 1, "xnut": 1, "wd": 1, "dtev": 1, "x": 1, "y": 1, "z": 1}


class Oracle:
    """One oracle state: module-like arrays over all levels."""

    def __init__(self, params):
        self.params = params.copy()
        self.h = lib().or_create(ctypes.byref(self.params))
        if not self.h:
            raise ValueError("or_create rejected the parameters")
        self.nall = lib().or_nall(self.h)
        self.n1 = params.img * params.jmg * params.kmg

    def __del__(self):
        h = getattr(self, "h", None)
        if h and _lib is not None:
            _lib.or_destroy(h)
            self.h = None

    def level_offset(self, n):
        return lib().or_level_offset(self.h, n)

    def load_case(self, case):
        lib().or_set_grid(self.h, _p(np.ascontiguousarray(case.x)), _p(np.ascontiguousarray(case.y)),
                          _p(np.ascontiguousarray(case.z)))
        lib().or_set_wd(self.h, _p(np.ascontiguousarray(case.wd)))
        q, qn, qnm1, xnut = case.state()
        self.set_state(q, qn, qnm1, xnut)
        if getattr(case, "overset", None) is not None:
            self.set_overset(case.overset)
        if getattr(case, "alm", None) is not None:
            self.set_turbines(case.alm)
        return self

    def set_turbines(self, alm):
        """Actuator-line tables (desalm.alm.AlmSetup); None removes them."""
        if alm is None:
            assert lib().or_set_turbines(self.h, 0, None, 0, None, 0.0, 0.0) == 0
            return
        ts, ps, keep = alm.as_ctypes()
        rc = lib().or_set_turbines(self.h, len(alm.turbines), ctypes.cast(ts, ctypes.c_void_p), len(alm.polars),
                                   ctypes.cast(ps, ctypes.c_void_p), alm.length_scale, alm.vel_infty)
        assert rc == 0
        del keep

    def set_hooks(self, hooks, wf_newton=False):
        """OR_HOOK_* flags; wf_newton: the reference's own u_tau Newton chain."""
        if lib().or_set_hooks(self.h, int(hooks), int(bool(wf_newton))) != 0:
            raise ValueError("bad hook flags")
        return self

    def set_overset(self, ov):
        """Donor table (desalm.cases.Overset) of the fine level."""
        ip = ctypes.POINTER(ctypes.c_int)
        h = np.ascontiguousarray(ov.hosts, dtype=np.int32)
        c = np.ascontiguousarray(ov.coef, dtype=np.float64)
        t = np.ascontiguousarray(ov.nodes, dtype=np.int32)
        rc = lib().or_set_overset(self.h, len(c), h.ctypes.data_as(ip), _p(c), t.ctypes.data_as(ip),
                                  int(ov.apply_mom))
        assert rc == 0, "or_set_overset rejected the table"

    def bintp(self, which):
        """One mg_bintp_mom (which = "mom") or mg_bintp_tur ("tur")."""
        lib().or_bintp(self.h, 0 if which == "mom" else 1)

    def set_state(self, q=None, qn=None, qnm1=None, xnut=None):
        c = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
        q, qn, qnm1, xnut = c(q), c(qn), c(qnm1), c(xnut)
        lib().or_set_state(self.h, _p(q), _p(qn), _p(qnm1), _p(xnut))

    def get_state(self):
        q = np.zeros((self.n1, 5)); qn = np.zeros((self.n1, 5)); qnm1 = np.zeros((self.n1, 5))
        xnut = np.zeros(self.n1)
        lib().or_get_state(self.h, _p(q), _p(qn), _p(qnm1), _p(xnut))
        return q, qn, qnm1, xnut

    def get(self, name):
        c = NCOMP[name]
        out = np.zeros((self.nall, c)) if c > 1 else np.zeros(self.nall)
        rc = lib().or_get_array(self.h, name.encode(), _p(out))
        assert rc == 0, name
        return out

    def set(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert a.size == NCOMP[name] * self.nall
        assert lib().or_set_array(self.h, name.encode(), _p(a)) == 0

    def mg_driver(self):
        itc = ctypes.c_int(0)
        hist = np.zeros((max(self.params.itm, 1), 5))
        rc = lib().or_mg_driver(self.h, ctypes.byref(itc), _p(hist))
        assert rc == 0
        return itc.value, hist[: itc.value]

    def solver_daf(self, n, recalc, calc_pk):
        assert lib().or_solver_daf(self.h, n, recalc, calc_pk) == 0

    def residual_only(self, n=1, calc_pk=0):
        assert lib().or_residual(self.h, n, calc_pk) == 0
        return self.get("rh")

    def dtau(self, n=1):
        """The local pseudo time step of level n (rhs_daf_dtau, or
        met_daf_dtau's formula per params.dtau_formula)."""
        ni, nj, nk = self.params.level_dims()[n - 1]
        out = np.zeros(ni * nj * nk)
        assert lib().or_dtau(self.h, n, _p(out)) == 0
        return out

    def des_eddy(self):
        assert lib().or_des_eddy(self.h) == 0

    def mg_nlevel(self, n):
        assert lib().or_mg_nlevel(self.h, n) == 0

    def mg_prolong(self, n):
        assert lib().or_mg_prolong(self.h, n) == 0

    def residual(self):
        r = np.zeros(5)
        lib().or_mg_residual(self.h, _p(r))
        return r


def gtsv_pivots(reset=True):
    """Row interchanges done by the oracle's SGTSV since the last reset."""
    return lib().or_gtsv_pivots(1 if reset else 0)


def gtsv(dl, d, du, b):
    """SGTSV algorithm in FP64 on copies; returns (x, info)."""
    n = d.shape[0]
    b = np.array(b, dtype=np.float64, order="F", copy=True)
    nrhs = 1 if b.ndim == 1 else b.shape[1]
    dl = np.ascontiguousarray(dl, dtype=np.float64).copy()
    d = np.ascontiguousarray(d, dtype=np.float64).copy()
    du = np.ascontiguousarray(du, dtype=np.float64).copy()
    bf = np.asfortranarray(b)
    info = lib().or_gtsv(n, nrhs, _p(dl) if n > 1 else _p(np.zeros(1)), _p(d),
                         _p(du) if n > 1 else _p(np.zeros(1)),
                         bf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), max(n, 1))
    return bf, info


# ---------------------------------------------------------------------------
# reference Fortran harness
# ---------------------------------------------------------------------------
def ref_available():
    return os.path.exists(REF_EXE)


def hooks_available():
    return os.path.exists(HOOKS_EXE)


def wf_root(u0, yp0, re0):
    """u_tau converged to rounding (the wall-function root)."""
    return lib().or_wf_root(u0, yp0, re0)


def wf_newt(x, u0, yp0, re0):
    """The reference's Newton for u_tau from start x; returns the result."""
    v = ctypes.c_double(x)
    if lib().or_wf_newt(ctypes.byref(v), u0, yp0, re0) != 0:
        raise RuntimeError("the reference's solve would stop the program here")
    return v.value


def group_mg_driver(oracles, table):
    """One physical step of several zones (Oracle objects, zone 1 first)
    coupled by a cross-zone donor table (desalm.overset.ZoneTable); returns
    (itc, norms history)."""
    ip = ctypes.POINTER(ctypes.c_int)
    arr = (ctypes.c_void_p * len(oracles))(*[o.h for o in oracles])
    if table is None:
        n, h, c, t, mom = 0, np.zeros((1, 4), np.int32), np.zeros((1, 3)), np.zeros((1, 4), np.int32), 0
    else:
        n, mom = len(table), int(table.apply_mom)
        h = np.ascontiguousarray(table.hosts, dtype=np.int32)
        c = np.ascontiguousarray(table.coef, dtype=np.float64)
        t = np.ascontiguousarray(table.nodes, dtype=np.int32)
    itc = ctypes.c_int(0)
    hist = np.zeros((max(oracles[0].params.itm, 1), 5))
    rc = lib().or_group_mg_driver(arr, len(oracles), n, h.ctypes.data_as(ip), _p(c), t.ctypes.data_as(ip), mom,
                                  ctypes.byref(itc), _p(hist))
    assert rc == 0, "or_group_mg_driver rejected the zones or the table"
    return itc.value, hist[: itc.value]


def group_bintp(oracles, table, which):
    """One cross-zone interpolation (which = "mom" or "tur") with a ZoneTable."""
    ip = ctypes.POINTER(ctypes.c_int)
    arr = (ctypes.c_void_p * len(oracles))(*[o.h for o in oracles])
    h = np.ascontiguousarray(table.hosts, dtype=np.int32)
    c = np.ascontiguousarray(table.coef, dtype=np.float64)
    t = np.ascontiguousarray(table.nodes, dtype=np.int32)
    rc = lib().or_group_bintp(arr, len(oracles), len(table), h.ctypes.data_as(ip), _p(c), t.ctypes.data_as(ip),
                              0 if which == "mom" else 1)
    assert rc == 0, "or_group_bintp rejected the table"


def _write_alm(f, td, alm):
    """turbine section of the harness input, and the polar files 'pNNN.txt'
    (angle Cl Cd per line, closed by the -999 sentinel, src/rhs_lines.F90:239)"""
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32).tofile(f)
    f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64).tofile(f)
    th = np.zeros((len(alm.turbines), 8), dtype=np.int32)
    rr = [alm.length_scale, alm.vel_infty]
    for t, T in enumerate(alm.turbines):
        th[t, :7] = (T.nblades, T.nelems, T.nangles, T.itur1, T.itur2, len(T.disk_nodes), len(T.ker_nodes))
    for T in alm.turbines:
        rr += [T.tur_rad, T.hub_rad, T.tsr, T.angle, T.dtheta_interp]
    i32(th)
    f64(rr)
    for T in alm.turbines:
        i32(T.node_indexes)
        f64(T.node_indexes_pos)
        f64(T.disk_dat)
        i32(T.polar_before)
        i32(T.polar_after)
        i32(T.disk_nodes)
        i32(T.ker_nodes)
        f64(T.ker_pos)
        f64(T.ker_dat)
    for k, (a, cl, cd) in enumerate(alm.polars):
        with open(os.path.join(td, "p%03d.txt" % (k + 1)), "w") as pf:
            for x, y, z in zip(a, cl, cd):
                pf.write("%.17g %.17g %.17g\n" % (x, y, z))
            pf.write("-999 0 0\n")


def _big_stack():
    """The reference's procedures keep their level-sized work arrays as
    automatic arrays on the stack (e.g. solver_daf's), which overflow the
    default 8 MiB stack beyond about 65^3 nodes on one rank: the child gets
    the largest stack the hard limit allows."""
    import resource
    soft, hard = resource.getrlimit(resource.RLIMIT_STACK)
    try:
        resource.setrlimit(resource.RLIMIT_STACK, (hard, hard))
    except (ValueError, OSError):
        pass


def run_reference(case, mode=1, nsteps=1, daf_level=1, recalc=0, calc_pk=0, iters=None, timeout=600, exe=None,
                  rotate=False):
    """Run the reference's own Fortran (compiled from its sources by
    oracle/ref/build_ref.sh) on `case`.  The reference is built with three
    multigrid levels (src/global.F90:17); a two-level case runs with iter(3)=0,
    which leaves level 2 untouched by the third level's zero correction.
    exe: another build of the same harness (SHIM_MGD_EXE, SHIM_DAF_EXE: the
    reference with the Fortran drop-in linked in place of one routine).
    rotate: before every mg_driver call the blade angles advance as the
    reference's physical time loop does (src/main.F90:42-49).
    Returns a dict of fine-level arrays and all-level metrics, plus the text
    of the 'conver' file under "conver" when the run wrote one and, for
    mode 1, the wall time of the mg_driver calls under "seconds"."""
    p = case.params
    if p.ng > 3:
        raise ValueError("the reference is compiled for 3 levels")
    hdr = np.zeros(40, dtype=np.int32)
    hdr[0] = mode
    hdr[2:5] = (p.img, p.jmg, p.kmg)
    hdr[5:8] = (p.icrs, p.jcrs, p.kcrs)
    it = list(iters) if iters is not None else [p.iter[n] if n < p.ng else 0 for n in range(3)]
    hdr[8:11] = it
    hdr[11] = p.itm
    hdr[12] = p.it_min
    hdr[13] = p.unsteady
    hdr[14] = p.des
    hdr[15] = p.itk
    hdr[16:22] = list(p.btype)
    hdr[22:28] = list(p.bdir)
    hdr[28:31] = (p.ifix, p.jfix, p.kfix)
    hdr[31] = nsteps
    hdr[32] = daf_level
    hdr[33] = recalc
    hdr[34] = calc_pk
    rhd = np.zeros(20)
    rhd[:10] = (p.er_min, p.eo_min, p.ren, p.beta, p.delti, p.cfl1, p.vnn1, p.cfl2, p.vnn2, p.pdiss_coef)
    rhd[10:13] = list(p.ep[0])
    rhd[13:16] = list(p.ep[1])
    n1 = p.img * p.jmg * p.kmg
    ov = getattr(case, "overset", None)
    hdr[35] = 0 if ov is None else len(ov.coef)
    alm = getattr(case, "alm", None)
    hdr[36] = 0 if alm is None else len(alm.turbines)
    hdr[37] = 1 if rotate else 0
    with tempfile.TemporaryDirectory() as td:
        fin = os.path.join(td, "in.bin")
        fout = os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            hdr.tofile(f)
            rhd.tofile(f)
            for a in (case.x, case.y, case.z, case.wd, case.q, case.qn, case.qnm1, case.xnut):
                np.ascontiguousarray(a, dtype=np.float64).tofile(f)
            if ov is not None:
                np.ascontiguousarray(ov.hosts, dtype=np.int32).tofile(f)
                np.ascontiguousarray(ov.coef, dtype=np.float64).tofile(f)
                np.ascontiguousarray(ov.nodes, dtype=np.int32).tofile(f)
            if alm is not None:
                _write_alm(f, td, alm)
        env = dict(os.environ)
        subprocess.run([exe or REF_EXE, fin, fout], check=True, cwd=td, timeout=timeout, env=env,
                       stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, preexec_fn=_big_stack)
        conver = os.path.join(td, "conver")
        text = open(conver).read() if os.path.exists(conver) else None
        timing = os.path.join(td, "timing")
        seconds = float(open(timing).read()) if os.path.exists(timing) else None
        with open(fout, "rb") as f:
            nall, n1r = np.fromfile(f, dtype=np.int32, count=2)
            assert n1r == n1
            out = {}
            out["q"] = np.fromfile(f, dtype=np.float64, count=5 * n1).reshape(n1, 5)
            out["qn"] = np.fromfile(f, dtype=np.float64, count=5 * n1).reshape(n1, 5)
            out["qnm1"] = np.fromfile(f, dtype=np.float64, count=5 * n1).reshape(n1, 5)
            out["xnut"] = np.fromfile(f, dtype=np.float64, count=n1)
            out["rh"] = np.fromfile(f, dtype=np.float64, count=4 * n1).reshape(n1, 4)
            out["csi"] = np.fromfile(f, dtype=np.float64, count=3 * nall).reshape(nall, 3)
            out["eta"] = np.fromfile(f, dtype=np.float64, count=3 * nall).reshape(nall, 3)
            out["zet"] = np.fromfile(f, dtype=np.float64, count=3 * nall).reshape(nall, 3)
            out["aj"] = np.fromfile(f, dtype=np.float64, count=nall)
            out["dtev"] = np.fromfile(f, dtype=np.float64, count=nall)
    out["conver"] = text
    out["seconds"] = seconds  # wall time of the mg_driver calls (mode 1)
    return out


def relerr(a, b):
    """Normwise relative difference max|a-b| / max|b| (0 if both are 0)."""
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.max(np.abs(b))
    num = np.max(np.abs(a - b))
    if den == 0.0:
        return float(num)
    return float(num / den)
