"""The two local pseudo time step formulas of the reference.

rhs_daf_dtau (src/rhs_daf_dtau.F90:51-83) is the one solver_daf calls; its
viscous limits are vnn1 / (1/Re + nu_t) / h^2 / g.  met_daf_dtau
(src/met_daf_dtau.F90:53-93) is the other, which no reference path calls; its
viscous limits are vnn1 Re / h^2 / g.  desalm_params.dtau_formula selects it.

The oracle's met form is pinned by what the mathematics fixes: with nu_t = 0
the two formulas are the same number (vnn1 Re = vnn1 / (1/Re)) up to
rounding, and the met form does not depend on nu_t at all while the live one
shrinks where nu_t is large.  The HIP path is then held to the oracle at
1e-10 over whole pseudo-iterations.
"""
import numpy as np
import pytest

from desalm import cases
from desalm.params import DTAU_LIVE, DTAU_MET
import pyoracle as po


def _case(xnut_value=None, formula=DTAU_LIVE):
    c = cases.wavy(13, 11, 9, ng=2, des=True, itm=2, iters=(1, 1))
    c.params.dtau_formula = formula
    if xnut_value is not None:
        c.xnut = np.full_like(c.xnut, xnut_value)
    return c


def _dtau(c):
    return po.Oracle(c.params).load_case(c).dtau(1)


def test_met_equals_live_without_eddy_viscosity():
    live = _dtau(_case(0.0, DTAU_LIVE))
    met = _dtau(_case(0.0, DTAU_MET))
    assert np.all(np.isfinite(live)) and np.all(live > 0)
    assert np.max(np.abs(met - live) / live) < 1e-14


def test_met_ignores_eddy_viscosity_live_does_not():
    met0 = _dtau(_case(0.0, DTAU_MET))
    met1 = _dtau(_case(50.0, DTAU_MET))
    assert np.array_equal(met0, met1)
    live0 = _dtau(_case(0.0, DTAU_LIVE))
    live1 = _dtau(_case(50.0, DTAU_LIVE))
    # nu_t only lowers the viscous limits: never a larger step, and a smaller
    # one where they bind
    assert np.all(live1 <= live0)
    assert np.any(live1 < 0.5 * live0)


@pytest.mark.gpu
def test_met_formula_gpu_matches_oracle(gpu):
    from desalm.lib import Solver

    c = cases.wavy(17, 17, 17, ng=2, des=True, itm=3, iters=(1, 2))
    c.params.dtau_formula = DTAU_MET
    o = po.Oracle(c.params).load_case(c)
    s = Solver(c.params).load_case(c)
    for _ in range(2):
        itc_o, ho = o.mg_driver()
        itc_g, hg = s.mg_driver()
        assert itc_o == itc_g
        assert np.max(np.abs(hg - ho)) < 1e-9
    qg, _, _, xg = s.get_state()
    qo, _, _, xo = o.get_state()
    for m in range(5):
        assert po.relerr(qg[:, m], qo[:, m]) < 1e-10, m
    assert po.relerr(xg, xo) < 1e-10
    s.close()
