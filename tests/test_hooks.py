"""The case hooks of mg_driver on the GPU (desalm_set_hooks, k_hooks.hip)
against the oracle (oracle/desalm_oracle.c), which is bit-exact with the
reference built with src/conservation.F90 and src/wf_mpi.F90
(tests/test_oracle_reference.py, tests/golden/wavy_mg_driver_hooks_sa.npz).

Tolerances:
- conservation alone: FP64 parity as every other mg_driver test, 1e-9 of
  each component's scale (FMA contraction on the device);
- wall functions: the GPU solves u_tau to rounding at every node, as the
  oracle does with wf_newton = 0; 1e-8 of the scale, since the flow angle
  at the wall comes out of acos, whose slope grows without bound near
  +-1 and turns a last-bit difference of its argument into a larger one of
  the angle; against the reference's
  Newton chain, which stops at |f| < 1e-4 and starts each node from the
  previous node's result, the velocities differ by that tolerance: the test
  bounds it at 2e-3 of each component's scale after two physical steps."""
import numpy as np
import pytest

import pyoracle as po
from desalm import cases
from desalm.lib import DesalmError, HOOK_CONSERVATION, HOOK_WALL_FUNCTIONS, Solver

pytestmark = pytest.mark.gpu

BOTH = HOOK_CONSERVATION | HOOK_WALL_FUNCTIONS


def case(ng=2, bt=(3, 3, 1, 1, 2, 2)):
    return cases.wavy(17, 17, 21, ng=ng, itm=3, des=True, btype=bt, iters=(1, 2, 3), seed=5, amp=0.05)


def run_gpu(c, hooks, steps=2):
    s = Solver(c.params).load_case(c).set_hooks(hooks)
    for _ in range(steps):
        s.mg_driver()
    out = s.get_state()
    s.close()
    return out


def run_oracle(c, hooks, wf_newton, steps=2):
    o = po.Oracle(c.params).load_case(c).set_hooks(hooks, wf_newton=wf_newton)
    for _ in range(steps):
        o.mg_driver()
    return o.get_state()


def close(a, b, tol):
    scale = np.maximum(np.abs(b).max(axis=0), 1e-300)
    return (np.abs(a - b).max(axis=0) <= tol * scale).all(), np.abs(a - b).max(axis=0) / scale


@pytest.mark.parametrize("hooks", [HOOK_CONSERVATION, HOOK_WALL_FUNCTIONS, BOTH])
@pytest.mark.parametrize("ng,bt", [(2, (3, 3, 1, 1, 2, 2)), (3, (0, 0, 1, 1, 2, 2))])
def test_hooks_match_oracle(gpu, hooks, ng, bt):
    c = case(ng, bt)
    q, _, _, xnut = run_gpu(c, hooks)
    qo, _, _, xnuto = run_oracle(c, hooks, wf_newton=False)
    ok, err = close(q, qo, 1e-9 if hooks == HOOK_CONSERVATION else 1e-8)
    assert ok, err
    assert np.all(np.isfinite(q))
    # the hooks did something: without them the state differs
    q0 = run_gpu(c, 0)[0]
    assert not np.allclose(q0, q, rtol=0, atol=1e-12)


def test_wall_functions_within_reference_newton_tolerance(gpu):
    c = case()
    q = run_gpu(c, BOTH)[0]
    qr = run_oracle(c, BOTH, wf_newton=True)[0]   # = the reference, bit for bit
    ok, err = close(q, qr, 2e-3)
    assert ok, err


def test_hooks_need_the_grid_and_one_zone(gpu):
    c = case()
    s = Solver(c.params)
    with pytest.raises(DesalmError, match="set_grid"):
        s.set_hooks(BOTH)
    s.close()
    s = Solver(c.params).load_case(c)
    with pytest.raises(DesalmError, match="unknown hook"):
        s.set_hooks(8)
    s.set_hooks(0)
    s.close()
    from desalm.lib import Group
    a = Solver(c.params).load_case(c).set_hooks(HOOK_WALL_FUNCTIONS)
    b = Solver(c.params).load_case(c)
    g = Group([a, b])
    with pytest.raises(DesalmError, match="single zone"):
        g.mg_driver()
    g.close()
    a.close()
    b.close()


def combined_case():
    """Actuator line, an inflow face, a characteristic outlet, j walls and
    both hooks on one curvilinear block: the features together."""
    from golden.make_golden import build
    return build("wavy", dict(ni=21, nj=17, nk=17, ng=2, itm=3, amp=0.02, des=True, iters=(1, 2, 3),
                              btype=(4, 5, 1, 1, 2, 2), alm=dict(i_rot=9, center=[0.5, 0.5], radius=0.3, hub=0.06)))


def test_everything_together_matches_oracle(gpu):
    c = combined_case()
    q, _, _, xnut = run_gpu(c, BOTH)
    qo, _, _, xnuto = run_oracle(c, BOTH, wf_newton=False)
    ok, err = close(q, qo, 1e-8)
    assert ok, err
    assert np.all(np.isfinite(q))
