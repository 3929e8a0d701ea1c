"""Failure detection of the driver (desalm_last_status, DESALM_ERR_NONFINITE,
the zero-pivot INFO) and the fallback of the FP32-DAF form and the warm
starts to the plain FP64 method.  Every condition is forced from the host
with NaN inputs, so every kernel still runs to its end.

- NaN in the velocity field of a laminar block: the pseudo-iteration's
  update norms are NaN, the driver stops after that iteration with
  DESALM_ERR_NONFINITE.
- The same NaN with the SA model on: the convection coefficients of the
  SA line solves through that node are NaN, so their pivots are, and the
  driver returns the INFO-like row (> 0) before any norm is read.
- NaN in one grid coordinate: the metrics around it, dtau and the pivots of
  the line solves through it are NaN: the driver returns the INFO-like row
  (> 0).  (A NaN eddy viscosity alone does not reach dtau: the minimum of
  its candidate steps skips a NaN one.)
- A learned warm start whose network weights are NaN: its start is NaN, the
  step is redone from the step's entry state with the plain method and the
  reference's start, and ends exactly where a plain step from the same state
  ends (bitwise), with one fallback counted.
"""
import numpy as np
import pytest

from desalm import cases
from desalm.lib import ERR_NONFINITE, DesalmError, Solver

pytestmark = pytest.mark.gpu


def small_case(des=True, **kw):
    return cases.wavy(17, 17, 17, ng=2, itm=4, des=des, iters=(1, 2), **kw)


def test_nan_velocity_stops_with_nonfinite_norms(gpu):
    c = small_case(des=False)
    s = Solver(c.params).load_case(c)
    q, qn, qnm1, xnut = c.state()
    q = q.copy()
    q[(c.n1 // 2), 1] = np.nan   # an interior node
    s.set_state(q, qn, qnm1, xnut)
    with pytest.raises(DesalmError, match=f"rc={ERR_NONFINITE}.*not finite after pseudo-iteration 1"):
        s.mg_driver()
    st = s.last_status()
    assert st["nonfinite_iteration"] == 1 and st["zero_pivot_row"] == 0 and st["fallbacks"] == 0
    s.close()


def test_nan_velocity_with_sa_reports_the_sa_pivot(gpu):
    c = small_case()
    s = Solver(c.params).load_case(c)
    q, qn, qnm1, xnut = c.state()
    q = q.copy()
    q[(c.n1 // 2), 1] = np.nan
    s.set_state(q, qn, qnm1, xnut)
    with pytest.raises(DesalmError, match=r"rc=[1-9][0-9]*: a line solve met a zero or non-finite pivot"):
        s.mg_driver()
    st = s.last_status()
    assert st["zero_pivot_row"] > 0 and st["nonfinite_iteration"] == 0
    s.close()


def test_nan_metrics_report_a_zero_pivot_row(gpu):
    c = small_case()
    x = c.x.copy()
    x[c.n1 // 2] = np.nan
    c.x = x
    s = Solver(c.params).load_case(c)
    with pytest.raises(DesalmError, match=r"rc=[1-9][0-9]*: a line solve met a zero or non-finite pivot"):
        s.mg_driver()
    st = s.last_status()
    assert st["zero_pivot_row"] > 0
    s.close()


def test_clean_run_reports_nothing(gpu):
    c = small_case()
    s = Solver(c.params).load_case(c)
    s.mg_driver()
    assert s.last_status() == dict(zero_pivot_row=0, nonfinite_iteration=0, diverged_iteration=0, fallbacks=0)
    s.close()


def test_learned_start_with_nan_weights_falls_back_to_plain(gpu):
    c = cases.alm_box(41, 25, 25, ng=2, itm=6, nelems=6, nangles=73, iters=(1, 2))
    s = Solver(c.params).load_case(c)
    s.ws_configure(hidden=16, samples_per_record=4096, capacity=1 << 16, seed=3)
    for _ in range(3):
        s.step()
        s.ws_record()
    s.ws_train(2, batch=1024)
    snap = s.snapshot()
    # the plain step from this state
    s.ws_mode(Solver.WS_NONE)
    s.step()
    plain = s.get_state()
    # the same step from a network with NaN weights
    s.restore(snap)
    w = s.ws_params()
    s.ws_params(set_to=np.full_like(w, np.nan))
    s.ws_mode(Solver.WS_LEARNED)
    itc, hist = s.step()
    st = s.last_status()
    assert st["fallbacks"] == 1 and st["nonfinite_iteration"] == 0
    got = s.get_state()
    for a, b in zip(got, plain):
        assert np.array_equal(a, b)
    assert np.all(np.isfinite(hist))
    s.close()
