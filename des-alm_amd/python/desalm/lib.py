"""ctypes binding of libdesalm.so (include/desalm.h).

This is the Python-side caller of the C ABI used by the tests and bench.py;
it adds nothing to the computation.  It raises if the HIP library has not been
built: there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

from .params import Params

PKG_ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
LIB_PATH = os.path.join(PKG_ROOT, "libdesalm.so")

FIELD = {"q": (0, 5), "qn": (1, 5), "qnm1": (2, 5), "qold_mg": (3, 5), "rh": (4, 4), "pk": (5, 4),
         "xnut": (6, 1), "csi": (7, 3), "eta": (8, 3), "zet": (9, 3), "aj": (10, 1), "dtev": (11, 1),
         "wd": (12, 1)}

_lib = None


ERR_NONFINITE = -2  # DESALM_ERR_NONFINITE


class DesalmError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise DesalmError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.desalm_init.argtypes = [ctypes.POINTER(Params), ctypes.c_int, ctypes.POINTER(vp)]
        L.desalm_nodes_all.argtypes = [vp]
        L.desalm_nodes_all.restype = ctypes.c_int64
        L.desalm_level_offset.argtypes = [vp, ctypes.c_int]
        L.desalm_level_offset.restype = ctypes.c_int64
        L.desalm_set_grid.argtypes = [vp, dp, dp, dp]
        L.desalm_set_metrics.argtypes = [vp, dp, dp, dp, dp]
        L.desalm_set_wall_distance.argtypes = [vp, dp, dp]
        L.desalm_set_state.argtypes = [vp, dp, dp, dp, dp]
        L.desalm_get_state.argtypes = [vp, dp, dp, dp, dp]
        L.desalm_mg_driver.argtypes = [vp, ip, dp]
        L.desalm_solver_daf.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int] + [dp] * 10
        L.desalm_des_eddy.argtypes = [vp] + [dp] * 10
        L.desalm_gtsv_batch.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp, dp, dp, ip]
        L.desalm_gtsv_batch_noalloc.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp, dp, dp, ip]
        L.desalm_get_level_field.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp]
        L.desalm_last_driver_ms.argtypes = [vp]
        L.desalm_last_driver_ms.restype = ctypes.c_double
        L.desalm_last_pivot_events.argtypes = [vp]
        L.desalm_last_pivot_events.restype = ctypes.c_int64
        L.desalm_last_status.argtypes = [vp, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                         ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int64)]
        L.desalm_last_status.restype = ctypes.c_int
        L.desalm_real_residual.argtypes = [vp, dp]
        L.desalm_real_residual.restype = ctypes.c_int
        L.desalm_probe.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32), dp]
        L.desalm_probe.restype = ctypes.c_int
        L.desalm_profile.argtypes = [vp, ctypes.c_int]
        L.desalm_profile.restype = ctypes.c_int
        L.desalm_profile_get.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int64),
                                         ctypes.POINTER(ctypes.c_double)]
        L.desalm_profile_get.restype = ctypes.c_int
        L.desalm_profile_name.argtypes = [ctypes.c_int]
        L.desalm_profile_name.restype = ctypes.c_char_p
        L.desalm_last_error.argtypes = [vp]
        L.desalm_last_error.restype = ctypes.c_char_p
        L.desalm_finalize.argtypes = [vp]
        i32p = ctypes.POINTER(ctypes.c_int32)
        L.desalm_set_overset.argtypes = [vp, ctypes.c_int64, i32p, dp, i32p, ctypes.c_int]
        L.desalm_bintp.argtypes = [vp, ctypes.c_int]
        L.desalm_group_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.POINTER(vp)]
        L.desalm_group_set_overset.argtypes = [vp, ctypes.c_int64, i32p, dp, i32p, ctypes.c_int]
        L.desalm_group_mg_driver.argtypes = [vp, ip, dp]
        L.desalm_group_destroy.argtypes = [vp]
        L.desalm_group_set_batched.argtypes = [vp, ctypes.c_int]
        L.desalm_set_turbines.argtypes = [vp, ctypes.c_int32, vp, ctypes.c_int32, vp, ctypes.c_double,
                                          ctypes.c_double]
        fp = ctypes.POINTER(ctypes.c_float)
        i64 = ctypes.c_int64
        L.desalm_set_convergence.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.desalm_set_hooks.argtypes = [vp, ctypes.c_int]
        L.desalm_set_coordinates.argtypes = [vp, dp, dp, dp]
        L.desalm_step.argtypes = [vp, ip, dp]
        L.desalm_ws_configure.argtypes = [vp, ctypes.c_int, i64, i64, ctypes.c_uint64]
        L.desalm_ws_mode.argtypes = [vp, ctypes.c_int]
        L.desalm_ws_record.argtypes = [vp]
        L.desalm_ws_train.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, dp]
        L.desalm_ws_param_count.argtypes = [vp]
        L.desalm_ws_params.argtypes = [vp, fp, fp]
        L.desalm_ws_stats.argtypes = [vp, fp, fp]
        L.desalm_ws_features.argtypes = [vp, i64, fp, fp]
        L.desalm_mlp_forward.argtypes = [vp, fp, i64, fp]
        L.desalm_mlp_loss_grad.argtypes = [vp, fp, fp, i64, fp, dp]
        L.desalm_ws_history.argtypes = [vp, dp, dp]
        L.desalm_set_tiles.argtypes = [vp, ip, ctypes.c_int]
        L.desalm_get_tiles.argtypes = [vp, ip, ip]
        L.desalm_autotune.argtypes = [vp, ctypes.c_int]
        L.desalm_set_rfg.argtypes = [vp, ctypes.c_int32] + [dp] * 9 + [ctypes.c_double]
        L.desalm_set_rfg.restype = ctypes.c_int
        L.desalm_device_bytes.argtypes = [vp]
        L.desalm_device_bytes.restype = ctypes.c_int64
        L.desalm_turbine_angle.argtypes = [vp, ctypes.c_int, dp, dp]
        for f in ("desalm_set_convergence", "desalm_step", "desalm_ws_configure", "desalm_ws_mode",
                  "desalm_ws_record", "desalm_ws_train", "desalm_ws_param_count", "desalm_ws_params",
                  "desalm_ws_stats", "desalm_ws_features", "desalm_mlp_forward", "desalm_mlp_loss_grad",
                  "desalm_ws_history", "desalm_turbine_angle", "desalm_set_tiles", "desalm_get_tiles",
                  "desalm_autotune"):
            getattr(L, f).restype = ctypes.c_int
        for f in ("desalm_init", "desalm_set_grid", "desalm_set_metrics", "desalm_set_wall_distance",
                  "desalm_set_state", "desalm_get_state", "desalm_mg_driver", "desalm_solver_daf",
                  "desalm_des_eddy", "desalm_gtsv_batch", "desalm_gtsv_batch_noalloc", "desalm_get_level_field", "desalm_finalize",
                  "desalm_set_overset", "desalm_bintp", "desalm_group_create", "desalm_group_set_overset",
                  "desalm_group_mg_driver", "desalm_group_destroy", "desalm_group_set_batched",
                  "desalm_set_turbines"):
            getattr(L, f).restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    if a.dtype != np.float64 or not a.flags.c_contiguous:
        raise TypeError("arrays must be C-contiguous float64")
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _f(a):
    if a is None:
        return None
    if a.dtype != np.float32 or not a.flags.c_contiguous:
        raise TypeError("arrays must be C-contiguous float32")
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _c(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


HOOK_CONSERVATION = 1
HOOK_WALL_FUNCTIONS = 2


class Solver:
    """One desalm context on one HIP device."""

    def __init__(self, params, device=0):
        self.params = params.copy()
        h = ctypes.c_void_p()
        rc = lib().desalm_init(ctypes.byref(self.params), device, ctypes.byref(h))
        if rc != 0:
            raise DesalmError("desalm_init: " + lib().desalm_last_error(None).decode())
        self.h = h
        self.nall = lib().desalm_nodes_all(h)
        self.n1 = params.img * params.jmg * params.kmg
        self.dims = params.level_dims()

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().desalm_finalize(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc != 0:
            raise DesalmError(f"{what}: rc={rc}: " + lib().desalm_last_error(self.h).decode())

    def load_case(self, case):
        self.set_grid(case.x, case.y, case.z)
        self.set_wall_distance(case.wd)
        self.set_state(*case.state())
        if getattr(case, "overset", None) is not None:
            self.set_overset(case.overset)
        if getattr(case, "alm", None) is not None:
            self.set_turbines(case.alm)
        return self

    def set_turbines(self, alm):
        """Actuator-line rotors (desalm.alm.AlmSetup); None turns the forcing off."""
        if alm is None:
            self._chk(lib().desalm_set_turbines(self.h, 0, None, 0, None, 0.0, 0.0), "desalm_set_turbines")
            self._nturb = 0
            return
        self._nturb = len(alm.turbines)
        ts, ps, keep = alm.as_ctypes()
        self._chk(lib().desalm_set_turbines(self.h, len(alm.turbines), ctypes.cast(ts, ctypes.c_void_p),
                                            len(alm.polars), ctypes.cast(ps, ctypes.c_void_p), alm.length_scale,
                                            alm.vel_infty), "desalm_set_turbines")
        del keep

    def set_overset(self, ov):
        """Donor table (desalm.cases.Overset) of the fine level; None removes it."""
        i32p = ctypes.POINTER(ctypes.c_int32)
        if ov is None:
            self._chk(lib().desalm_set_overset(self.h, 0, None, None, None, 0), "desalm_set_overset")
            return
        h = np.ascontiguousarray(ov.hosts, dtype=np.int32)
        c = np.ascontiguousarray(ov.coef, dtype=np.float64)
        t = np.ascontiguousarray(ov.nodes, dtype=np.int32)
        self._chk(lib().desalm_set_overset(self.h, len(c), h.ctypes.data_as(i32p), _p(c), t.ctypes.data_as(i32p),
                                           int(ov.apply_mom)), "desalm_set_overset")

    def bintp(self, which):
        """One mg_bintp_mom (which = "mom") or mg_bintp_tur ("tur") on the device state."""
        self._chk(lib().desalm_bintp(self.h, 0 if which == "mom" else 1), "desalm_bintp")

    def set_grid(self, x, y, z):
        x, y, z = _c(x), _c(y), _c(z)
        self._chk(lib().desalm_set_grid(self.h, _p(x), _p(y), _p(z)), "desalm_set_grid")

    def set_metrics(self, csi, eta, zet, aj):
        a = [_c(v) for v in (csi, eta, zet, aj)]
        self._chk(lib().desalm_set_metrics(self.h, *[_p(v) for v in a]), "desalm_set_metrics")

    def set_wall_distance(self, wd, dtev=None):
        wd, dtev = _c(wd), _c(dtev)
        self._chk(lib().desalm_set_wall_distance(self.h, _p(wd), _p(dtev)), "desalm_set_wall_distance")

    def set_state(self, q=None, qn=None, qnm1=None, xnut=None):
        a = [_c(v) for v in (q, qn, qnm1, xnut)]
        self._chk(lib().desalm_set_state(self.h, *[_p(v) for v in a]), "desalm_set_state")

    def get_state(self):
        q = np.zeros((self.n1, 5)); qn = np.zeros((self.n1, 5)); qnm1 = np.zeros((self.n1, 5))
        xnut = np.zeros(self.n1)
        self._chk(lib().desalm_get_state(self.h, _p(q), _p(qn), _p(qnm1), _p(xnut)), "desalm_get_state")
        return q, qn, qnm1, xnut

    def mg_driver(self):
        itc = ctypes.c_int(0)
        hist = np.zeros((max(self.params.itm, 1), 5))
        self._chk(lib().desalm_mg_driver(self.h, ctypes.byref(itc), _p(hist)), "desalm_mg_driver")
        return itc.value, hist[: itc.value]

    def set_rfg(self, modes, tables, time0=0.0):
        """Random-flow-generator inflow: modes = (k, omega, p, q) and tables =
        (um, m, c, ts, ls) as built by desalm.rfg; None removes it."""
        if modes is None:
            self._chk(lib().desalm_set_rfg(self.h, 0, *([None] * 9), 0.0), "desalm_set_rfg")
            return
        k, omega, p, q = (_c(a).ravel() for a in modes)
        um, m, c, ts, ls = (_c(a).ravel() for a in tables)
        self._keep_rfg = (k, omega, p, q, um, m, c, ts, ls)
        self._chk(lib().desalm_set_rfg(self.h, len(omega), *[_p(a) for a in self._keep_rfg], time0),
                  "desalm_set_rfg")

    @property
    def device_bytes(self):
        """device memory held by the context (fields, scratch, rotors, warm start)"""
        return lib().desalm_device_bytes(self.h)

    def set_tiles(self, kc_visc=None, kc_sa=0):
        """k planes per workgroup of the LDS-tiled kernels (0: default)."""
        arr = (ctypes.c_int * self.params.ng)(*([0] * self.params.ng if kc_visc is None else kc_visc))
        self._chk(lib().desalm_set_tiles(self.h, arr, kc_sa), "desalm_set_tiles")

    def get_tiles(self):
        arr = (ctypes.c_int * self.params.ng)()
        sa = ctypes.c_int(0)
        self._chk(lib().desalm_get_tiles(self.h, arr, ctypes.byref(sa)), "desalm_get_tiles")
        return list(arr), sa.value

    def autotune(self, reps=5):
        """Time every launch shape of the LDS-tiled kernels on the current
        state and keep the fastest; returns get_tiles()."""
        self._chk(lib().desalm_autotune(self.h, reps), "desalm_autotune")
        return self.get_tiles()

    def set_coordinates(self, x, y, z):
        """Fine-level coordinates only (metrics from set_metrics stay)."""
        self._chk(lib().desalm_set_coordinates(self.h, _p(x), _p(y), _p(z)), "desalm_set_coordinates")

    def set_hooks(self, hooks):
        """HOOK_CONSERVATION | HOOK_WALL_FUNCTIONS (the reference's mg_driver
        calls both on every run; 0 turns them off)."""
        self._chk(lib().desalm_set_hooks(self.h, int(hooks)), "desalm_set_hooks")
        return self

    def set_convergence(self, itm, it_min, er_min, eo_min):
        self._chk(lib().desalm_set_convergence(self.h, itm, it_min, er_min, eo_min), "desalm_set_convergence")
        self.params.itm, self.params.it_min = itm, it_min
        self.params.er_min, self.params.eo_min = er_min, eo_min

    def step(self):
        """One physical time step: blade angles advance, then mg_driver from
        the selected warm start.  Returns (itc, log10 norm history)."""
        itc = ctypes.c_int(0)
        hist = np.zeros((max(self.params.itm, 1), 5))
        self._chk(lib().desalm_step(self.h, ctypes.byref(itc), _p(hist)), "desalm_step")
        return itc.value, hist[: itc.value]

    # ---- learned warm start ----
    WS_NONE, WS_EXTRAP, WS_LEARNED = 0, 1, 2

    def ws_configure(self, hidden=32, samples_per_record=65536, capacity=1 << 22, seed=20260505):
        self._chk(lib().desalm_ws_configure(self.h, hidden, samples_per_record, capacity, seed),
                  "desalm_ws_configure")
        self._ws = True

    def ws_mode(self, mode):
        self._chk(lib().desalm_ws_mode(self.h, mode), "desalm_ws_mode")

    def ws_record(self):
        self._chk(lib().desalm_ws_record(self.h), "desalm_ws_record")

    def ws_train(self, iters, batch=16384, lr=1e-3):
        loss = np.zeros(max(iters, 1))
        self._chk(lib().desalm_ws_train(self.h, iters, batch, lr, _p(loss)), "desalm_ws_train")
        return loss[:iters]

    def ws_params(self, set_to=None):
        n = lib().desalm_ws_param_count(self.h)
        out = np.zeros(n, dtype=np.float32)
        src = None if set_to is None else np.ascontiguousarray(set_to, dtype=np.float32)
        self._chk(lib().desalm_ws_params(self.h, _f(out), _f(src)), "desalm_ws_params")
        return out

    def ws_stats(self, set_to=None):
        out = np.zeros(2 * 95 + 2 * 5, dtype=np.float32)
        src = None if set_to is None else np.ascontiguousarray(set_to, dtype=np.float32)
        self._chk(lib().desalm_ws_stats(self.h, _f(out), _f(src)), "desalm_ws_stats")
        return out

    def ws_history(self, set_to=None):
        """q^(n-2) (n1, 5); set_to restores it."""
        out = np.zeros((self.n1, 5))
        self._chk(lib().desalm_ws_history(self.h, _p(out), _p(_c(set_to))), "desalm_ws_history")
        return out

    def turbine_angle(self, t=0, set_to=None):
        a = ctypes.c_double(0.0)
        b = None if set_to is None else ctypes.byref(ctypes.c_double(set_to))
        self._chk(lib().desalm_turbine_angle(self.h, t, ctypes.byref(a), b), "desalm_turbine_angle")
        return a.value

    def snapshot(self):
        """Everything a physical step reads: state, q^(n-2) (with a warm
        start) and blade angles."""
        st = self.get_state()
        hist = self.ws_history() if getattr(self, "_ws", False) else None
        ang = [self.turbine_angle(t) for t in range(getattr(self, "_nturb", 0))]
        return st, hist, ang

    def restore(self, snap):
        (q, qn, qnm1, xnut), hist, ang = snap
        self.set_state(q, qn, qnm1, xnut)
        if hist is not None:
            self.ws_history(hist)
        for t, a in enumerate(ang):
            self.turbine_angle(t, a)

    def ws_features(self, m):
        x = np.zeros((m, 95), dtype=np.float32)
        t = np.zeros((m, 15), dtype=np.float32)
        self._chk(lib().desalm_ws_features(self.h, m, _f(x), _f(t)), "desalm_ws_features")
        return x, t

    def mlp_forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((x.shape[0], 10), dtype=np.float32)
        self._chk(lib().desalm_mlp_forward(self.h, _f(x), x.shape[0], _f(y)), "desalm_mlp_forward")
        return y

    def mlp_loss_grad(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(lib().desalm_ws_param_count(self.h), dtype=np.float32)
        loss = ctypes.c_double(0.0)
        self._chk(lib().desalm_mlp_loss_grad(self.h, _f(x), _f(t), x.shape[0], _f(g), ctypes.byref(loss)),
                  "desalm_mlp_loss_grad")
        return loss.value, g

    def solver_daf(self, n, recalc, calc_pk, q, qn, qnm1, csi, eta, zet, aj, xnut, pk, rh):
        """In-place on q, pk, rh (float64 C-contiguous arrays of level n)."""
        args = [q, _c(qn), _c(qnm1), _c(csi), _c(eta), _c(zet), _c(aj), _c(xnut), pk, rh]
        self._chk(lib().desalm_solver_daf(self.h, n, recalc, calc_pk, *[_p(a) for a in args]), "desalm_solver_daf")

    def des_eddy(self, q, qn5, qnm1_5, csi, eta, zet, aj, wd, dtev, xnut):
        """In-place on q (n1,5) and xnut (n1)."""
        args = [q, _c(qn5), _c(qnm1_5), _c(csi), _c(eta), _c(zet), _c(aj), _c(wd), _c(dtev), xnut]
        self._chk(lib().desalm_des_eddy(self.h, *[_p(a) for a in args]), "desalm_des_eddy")

    def level_field(self, name, n=1):
        which, nc = FIELD[name]
        i, j, k = self.dims[n - 1]
        out = np.zeros((i * j * k, nc)) if nc > 1 else np.zeros(i * j * k)
        self._chk(lib().desalm_get_level_field(self.h, which, n, _p(out)), "desalm_get_level_field")
        return out

    def profile(self, enable=True):
        self._chk(lib().desalm_profile(self.h, 1 if enable else 0), "desalm_profile")

    def profile_stats(self):
        """{(kernel_name, level): (launches, total_ms)} of the recorded launches."""
        out = {}
        kid = 0
        while True:
            nm = lib().desalm_profile_name(kid)
            if nm is None:
                break
            for n in range(1, self.params.ng + 1):
                cnt = ctypes.c_int64(0)
                ms = ctypes.c_double(0.0)
                lib().desalm_profile_get(self.h, kid, n, ctypes.byref(cnt), ctypes.byref(ms))
                if cnt.value:
                    out[(nm.decode(), n)] = (cnt.value, ms.value)
            kid += 1
        return out

    @property
    def last_ms(self):
        return lib().desalm_last_driver_ms(self.h)

    @property
    def pivot_events(self):
        return lib().desalm_last_pivot_events(self.h)

    def real_residual(self):
        """log10 sqrt(mean over the fine interior of (q - qn)^2) per component:
        the 'conver_real' line of the reference (src/main.F90:191-300)."""
        out = np.zeros(5)
        self._chk(lib().desalm_real_residual(self.h, _p(out)), "desalm_real_residual")
        return out

    def probe(self, ijk):
        """q(1:5) and xnut (rows of 6) at 1-based fine-level points (n, 3): the
        'history' values of the reference's monitor points."""
        pts = np.ascontiguousarray(np.asarray(ijk, dtype=np.int32).reshape(-1, 3))
        out = np.zeros((len(pts), 6))
        self._chk(lib().desalm_probe(self.h, len(pts), pts.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                     _p(out)), "desalm_probe")
        return out

    def last_status(self):
        """Failure detection of the last driver call: dict of zero_pivot_row,
        nonfinite_iteration, diverged_iteration (0 = none) and the number of
        fallbacks to the plain FP64 method so far."""
        z, nf, dv = ctypes.c_int32(0), ctypes.c_int32(0), ctypes.c_int32(0)
        fb = ctypes.c_int64(0)
        self._chk(lib().desalm_last_status(self.h, ctypes.byref(z), ctypes.byref(nf), ctypes.byref(dv),
                                           ctypes.byref(fb)), "desalm_last_status")
        return dict(zero_pivot_row=z.value, nonfinite_iteration=nf.value, diverged_iteration=dv.value,
                    fallbacks=fb.value)


class Group:
    """Zones (Solver contexts, zone 1 first) coupled by a cross-zone donor
    table; desalm_group_* of include/desalm.h."""

    def __init__(self, solvers):
        self.solvers = list(solvers)
        arr = (ctypes.c_void_p * len(self.solvers))(*[sv.h.value for sv in self.solvers])
        self.h = ctypes.c_void_p()
        rc = lib().desalm_group_create(arr, len(self.solvers), ctypes.byref(self.h))
        if rc != 0:
            raise DesalmError(f"desalm_group_create: rc={rc}: " + lib().desalm_last_error(None).decode())

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().desalm_group_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_overset(self, table):
        """table: desalm.overset.ZoneTable (hosts/nodes (n, 4) i, j, k, zone; coef (n, 3))."""
        i32p = ctypes.POINTER(ctypes.c_int32)
        h = np.ascontiguousarray(table.hosts, dtype=np.int32)
        c = np.ascontiguousarray(table.coef, dtype=np.float64)
        t = np.ascontiguousarray(table.nodes, dtype=np.int32)
        rc = lib().desalm_group_set_overset(self.h, len(c), h.ctypes.data_as(i32p), _p(c), t.ctypes.data_as(i32p),
                                            int(table.apply_mom))
        if rc != 0:
            raise DesalmError(f"desalm_group_set_overset: rc={rc}: " +
                              lib().desalm_last_error(self.solvers[0].h).decode())

    def set_batched(self, on=True):
        """One launch per kernel over all zones (the default), or zone by zone."""
        rc = lib().desalm_group_set_batched(self.h, 1 if on else 0)
        if rc != 0:
            raise DesalmError(f"desalm_group_set_batched: rc={rc}")

    def mg_driver(self):
        itm = self.solvers[0].params.itm
        itc = ctypes.c_int(0)
        hist = np.zeros((max(itm, 1), 5))
        rc = lib().desalm_group_mg_driver(self.h, ctypes.byref(itc), _p(hist))
        if rc != 0:
            raise DesalmError(f"desalm_group_mg_driver: rc={rc}: " +
                              lib().desalm_last_error(self.solvers[0].h).decode())
        return itc.value, hist[: itc.value]


def gtsv_batch(dl, d, du, b, noalloc=False):
    """Batched SGTSV on the device.  Arrays (nsys, n); returns (x, info).
    noalloc: through desalm_gtsv_batch_noalloc (fixed device staging, at most
    DESALM_GTSV_NOALLOC_MAX = n * nsys entries)."""
    d = np.array(d, dtype=np.float64, copy=True, order="C")
    nsys, n = d.shape
    dl = np.array(dl, dtype=np.float64, copy=True, order="C").reshape(nsys, n)
    du = np.array(du, dtype=np.float64, copy=True, order="C").reshape(nsys, n)
    b = np.array(b, dtype=np.float64, copy=True, order="C").reshape(nsys, n)
    info = np.zeros(nsys, dtype=np.int32)
    f = lib().desalm_gtsv_batch_noalloc if noalloc else lib().desalm_gtsv_batch
    rc = f(n, nsys, _p(dl), _p(d), _p(du), _p(b), info.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    if rc != 0:
        raise DesalmError("desalm_gtsv_batch: " + lib().desalm_last_error(None).decode())
    return b, info
