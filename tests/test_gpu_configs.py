"""The five workloads of the benchmark on the GPU against the oracle (the C
restatement, bit-exact with the reference's Fortran), at sizes the oracle
finishes in seconds, and the full sizes through properties that do not need
it.

- Channel (SA, three levels): 33^3 with the reference's periodic-copy
  faces and 65^3 periodic (wrap) with the body force, in FP64, 1e-10
  normwise per field as every parity test; the FP32-DAF mixed form
  within 1e-6 of the FP64 oracle, and over a whole 300-iteration physical
  step of a curvilinear block at the same convergence as FP64 and within
  1e-6 of its end state.
- Actuator-line box (SA-DES, inflow, rotor of three blades): two physical
  steps of desalm_step, blades advancing, against the oracle stepped the same
  way: 1e-10.
- Lid-driven cavity (laminar, steady, two levels): the full 200
  pseudo-iterations: 1e-10 and identical norm histories to 1e-9.
- Overset: six hill blocks of uneven sizes coupled by the project's donor
  search, zone group against the oracle's group driver: 1e-10, with one
  launch per kernel over all blocks and zone by zone.
- Full sizes (129^3 channel, 193x129x129 and 257x193x193 actuator-line
  boxes): finite state, no row where SGTSV would interchange, and the HIP
  graph replay of the pseudo-iteration equal to eager launches.
"""
import math

import numpy as np
import pytest

from desalm import cases, overset
from desalm.lib import Group, Solver
from desalm.params import PREC_MIXED
import pyoracle as po

TOL = 1e-10
pytestmark = pytest.mark.gpu


def rel(a, b):
    return po.relerr(a, b)


def check_fields(s, o, tol=TOL):
    qg, qng, qnm1g, xg = s.get_state()
    qo, qno, qnm1o, xo = o.get_state()
    for m in range(5):
        assert rel(qg[:, m], qo[:, m]) < tol, m
    assert rel(xg, xo) < tol
    return qg, qo


@pytest.mark.parametrize("n,periodic", [(33, "copy"), (65, "wrap")])
def test_channel_fp64_matches_oracle(gpu, n, periodic):
    c = cases.channel(n, n, n, ng=3, itm=3, iters=(1, 2, 3), periodic=periodic)
    o = po.Oracle(c.params).load_case(c)
    s = Solver(c.params).load_case(c)
    for _ in range(2):
        itc_o, ho = o.mg_driver()
        itc_g, hg = s.mg_driver()
        assert itc_o == itc_g
        assert np.max(np.abs(hg - ho)) < 1e-9
    check_fields(s, o)
    assert s.pivot_events == 0


def test_channel_mixed_within_1e6_of_fp64_oracle(gpu):
    c = cases.channel(33, 33, 33, ng=3, itm=6, iters=(1, 2, 3))
    o = po.Oracle(c.params).load_case(c)
    c.params.precision = PREC_MIXED
    s = Solver(c.params).load_case(c)
    for _ in range(2):
        o.mg_driver()
        s.mg_driver()
    qg, qo = s.get_state()[0], o.get_state()[0]
    errs = [rel(qg[:, m], qo[:, m]) for m in range(5)]
    assert max(errs) < 1e-6, errs
    assert max(errs) > 0.0  # the FP32 sweeps ran


def test_mixed_converges_to_the_fp64_solution(gpu):
    """The FP32-DAF form over a whole physical step: 300 pseudo-iterations
    of the 33^3 curvilinear block (no early exit), FP64 and mixed from the
    same state.  Both drop the u, v, w update norms by the same 1.5 decades
    within a few iterations of each other, and the mixed state ends within
    1e-6 (normwise, per field) of the FP64 state."""
    def run(prec):
        c = cases.wavy(33, 33, 33, ng=3, itm=300, iters=(1, 2, 3), delti=0.02)
        c.params.er_min = -30.0
        c.params.eo_min = -30.0
        c.params.it_min = 0
        c.params.precision = prec
        s = Solver(c.params).load_case(c)
        itc, h = s.mg_driver()
        q = s.get_state()[0]
        s.close()
        m = h[:, 1:4].max(axis=1)
        return itc, m - m[0], q

    itc64, d64, q64 = run(0)
    itcm, dm, qm = run(PREC_MIXED)
    assert itc64 == itcm == 300
    target = -1.5
    assert d64[-1] < target and dm[-1] < target, (d64[-1], dm[-1])
    first64 = int(np.argmax(d64 < target))
    firstm = int(np.argmax(dm < target))
    assert abs(firstm - first64) <= max(3, first64 // 20), (first64, firstm)
    errs = [rel(qm[:, m], q64[:, m]) for m in range(4)]
    assert max(errs) < 1e-6, errs


def step_angles(alm, delti):
    """The blade-angle update of a physical step (src/main.F90:42-49)."""
    for T in alm.turbines:
        if T.angle > 2 * math.pi:
            T.angle -= 2 * math.pi
        T.angle += T.tsr * delti * alm.length_scale / T.tur_rad


def test_actuator_line_box_two_steps_match_oracle(gpu):
    c = cases.alm_box(49, 33, 33, ng=3, itm=4, nangles=181, nelems=12)
    s = Solver(c.params).load_case(c)
    o = po.Oracle(c.params).load_case(c)
    for _ in range(2):
        itc_g, hg = s.step()
        step_angles(c.alm, c.params.delti)
        o.set_turbines(c.alm)
        itc_o, ho = o.mg_driver()
        assert itc_g == itc_o
        assert np.max(np.abs(hg - ho)) < 1e-9
    assert abs(s.turbine_angle(0) - c.alm.turbines[0].angle) == 0.0
    check_fields(s, o)
    assert s.pivot_events == 0


def test_cavity_full_200_iterations_match_oracle(gpu):
    c = cases.cavity(33, ng=2, itm=200)
    c.params.it_min = 200
    o = po.Oracle(c.params).load_case(c)
    s = Solver(c.params).load_case(c)
    itc_o, ho = o.mg_driver()
    itc_g, hg = s.mg_driver()
    assert itc_o == itc_g == 200
    fin = np.isfinite(ho)
    assert np.array_equal(fin, np.isfinite(hg))
    assert np.max(np.abs(hg[fin] - ho[fin])) < 1e-9
    check_fields(s, o)
    assert s.pivot_events == 0


# six blocks of uneven sizes (the overset workload's proportions at about a
# fifth of its node count per direction), odd sizes for two levels
HILL_BLOCKS = [(21, 15, 11), (23, 17, 13), (27, 15, 15), (29, 21, 15), (33, 21, 17), (17, 11, 9)]


@pytest.mark.parametrize("batched", [True, False], ids=["one_launch", "zone_by_zone"])
def test_six_hill_blocks_match_oracle_group(gpu, batched):
    cs, tab = overset.hill_blocks(HILL_BLOCKS, [0.15, 0.32, 0.5, 0.68, 0.86], itm=3, amp=0.2, ren=10595.0)
    ors = [po.Oracle(c.params).load_case(c) for c in cs]
    svs = [Solver(c.params).load_case(c) for c in cs]
    g = Group(svs)
    g.set_batched(batched)
    g.set_overset(tab)
    for _ in range(2):
        itc_o, ho = po.group_mg_driver(ors, tab)
        itc_g, hg = g.mg_driver()
        assert itc_o == itc_g
        assert np.max(np.abs(hg - ho)) < 1e-9
    for o, s in zip(ors, svs):
        check_fields(s, o)
        assert s.pivot_events == 0
    g.close()


def graph_equals_eager(c, tiles=None):
    """One physical step replayed from a HIP graph and the same step launched
    eagerly (the event profiler turns the graph off) from the same state."""
    s = Solver(c.params).load_case(c)
    if tiles is not None:
        s.set_tiles(*tiles)
    snap = s.snapshot()
    s.mg_driver()
    a = s.get_state()
    piv = s.pivot_events
    s.restore(snap)
    s.profile(True)
    s.mg_driver()
    s.profile(False)
    b = s.get_state()
    s.close()
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    assert all(np.all(np.isfinite(x)) for x in a)
    assert piv == 0


def test_full_size_channel_129(gpu):
    graph_equals_eager(cases.channel(129, 129, 129, ng=3, itm=3, iters=(1, 2, 3)))


def test_full_size_actuator_line_box(gpu):
    graph_equals_eager(cases.alm_box(193, 129, 129, ng=3, itm=3))


def test_full_size_large_box(gpu):
    c = cases.alm_box(257, 193, 193, ng=4, itm=2, iters=(1, 2, 3, 4))
    c.params.precision = PREC_MIXED
    graph_equals_eager(c)
