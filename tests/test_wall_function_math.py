"""The wall-function friction velocity of the oracle (oracle/desalm_oracle.c,
wf_root and the reference's Newton chain wf_newt, after src/wf_mpi.F90:586-720)
pinned by the equation it solves:
  f(u) = (ln(Re y u) / 0.41 + 5.45) u - u0   if Re y u > 11.3,
  f(u) = Re y u^2 - u0                        otherwise,
  f = 0 where |u Re y - 11.3| < 1e-4.
wf_root must be a root to rounding on the log branch, the closed form
sqrt(u0 / (Re y)) on the quadratic branch and the jump point where f changes
sign across the jump; the reference's Newton must stop with |f| < 1e-4 (its
TOLF) or at its start when that already gives |f| < 1e-6."""
import math

import numpy as np

import pyoracle as po


def f(u, u0, y, re):
    if abs(u * re * y - 11.3) < 1e-4:
        return 0.0
    if re * y * u > 11.3:
        return (math.log(re * y * u) / 0.41 + 5.45) * u - u0
    return re * u * u * y - u0


def cases_(n, seed):
    rng = np.random.default_rng(seed)
    for _ in range(n):
        yield rng.uniform(1e-4, 2.0), 10 ** rng.uniform(-4, -1), 10 ** rng.uniform(2, 6)


def test_root_solves_the_equation():
    log_branch = quad_branch = jump = 0
    for u0, y, re in cases_(3000, 1):
        u = po.wf_root(u0, y, re)
        c = re * y
        if c * math.sqrt(u0 / c) <= 11.3:
            assert u == math.sqrt(u0 / c)
            quad_branch += 1
        elif u == 11.3 / c:
            # the jump: f just below is negative, just above non-negative
            assert c * u * u - u0 < 0 and (math.log(11.3) / 0.41 + 5.45) * u - u0 >= -1e-12 * u0
            jump += 1
        else:
            assert abs(f(u, u0, y, re)) <= 8 * np.finfo(float).eps * u0, (u0, y, re)
            # and f changes sign within 1e-12 of it
            assert f(u * (1 - 1e-12), u0, y, re) < 0 < f(u * (1 + 1e-12), u0, y, re)
            log_branch += 1
    assert log_branch > 100 and quad_branch > 100


def test_root_special_values():
    assert po.wf_root(0.0, 0.01, 1e4) == 0.0
    assert math.isnan(po.wf_root(float("nan"), 0.01, 1e4))
    # u0 exactly at the top of the quadratic branch
    re, y = 1e4, 1e-3
    u = 0.5
    u0 = re * y * u * u
    assert abs(po.wf_root(u0, y, re) - math.sqrt(u0 / (re * y))) == 0.0


def test_reference_newton_stops_inside_its_tolerance():
    for u0, y, re in cases_(2000, 2):
        x = po.wf_newt(0.03, u0, y, re)
        fx = f(x, u0, y, re)
        assert abs(fx) < 1e-4 or x == 0.03 and abs(f(0.03, u0, y, re)) < 1e-6
        # and it is the root's neighbour: |x - root| bounded by |f(x)| / min slope
        r = po.wf_root(u0, y, re)
        c = re * y
        lo = min(x, r)
        slope = (math.log(c * lo) / 0.41 + 5.45) if c * lo > 11.3 else 2 * c * lo
        if slope > 1.0 and abs(c * r - 11.3) > 1e-3:
            assert abs(x - r) <= 1e-4 / slope * 1.01 + 1e-12


def test_a_start_that_already_solves_is_kept():
    u0, y, re = 0.7, 1e-2, 1e4
    r = po.wf_root(u0, y, re)
    assert po.wf_newt(r, u0, y, re) == r
