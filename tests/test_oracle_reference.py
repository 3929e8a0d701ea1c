"""The CPU oracle against the reference Fortran itself.

Golden fixtures (tests/golden/*.npz) hold outputs of the reference's own
sources compiled in FP64 (oracle/ref/build_ref.sh, tests/golden/
make_golden.py).  The oracle must reproduce them bit for bit: it follows the
reference's loops and operand order, and gcc on x86-64 forms no FMAs.
When the reference harness is present (this container) the comparison is
also repeated live on a fresh input.
"""
import glob
import json
import os

import numpy as np
import pytest

from desalm import cases
import pyoracle as po
from golden.make_golden import build, conver_table, fingerprint

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))


def _case(meta):
    kw = {k: tuple(v) if isinstance(v, list) else v for k, v in meta["kwargs"].items()}
    return build(meta["kind"], kw)


def _run_oracle(case, opts):
    """Oracle run of a harness mode; returns the oracle and, for mg_driver,
    the per-iteration rows (iteration, 5 log10 norms) of every step."""
    o = po.Oracle(case.params).load_case(case)
    if opts.get("hooks"):
        o.set_hooks(opts["hooks"], wf_newton=True)
    mode = opts["mode"]
    rows = []
    if mode == 1:
        for _ in range(opts.get("nsteps", 1)):
            itc, hist = o.mg_driver()
            rows += [[it + 1] + list(hist[it]) for it in range(itc)]
    elif mode == 2:
        o.solver_daf(1, opts["recalc"], opts["calc_pk"])
    elif mode == 3:
        o.des_eddy()
    elif mode == 5:
        o.bintp("mom")
    elif mode == 6:
        o.bintp("tur")
    return o, np.array(rows)


def check_conver(rows, conver):
    """Oracle norms against the reference's 'conver' lines, printed g13.6
    (six significant digits); -Inf where a component does not move."""
    assert rows.shape == conver.shape
    assert np.array_equal(rows[:, 0], conver[:, 0])
    a, b = rows[:, 1:], conver[:, 1:]
    fin = np.isfinite(b)
    assert np.array_equal(np.isfinite(a), fin)
    assert np.array_equal(a[~fin], b[~fin])
    assert np.all(np.abs(a[fin] - b[fin]) <= 1e-5 * np.abs(b[fin]) + 1e-9)


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_oracle_matches_reference_golden(path):
    z = np.load(path, allow_pickle=False)
    meta = json.loads(str(z["meta"]))
    case = _case(meta)
    # the fixture was produced from exactly these inputs
    assert fingerprint(case) == meta["inputs_sha256"], "case builder changed; regenerate the fixtures"
    o, rows = _run_oracle(case, meta["options"])
    n1 = case.n1
    if meta["options"]["mode"] == 4:
        for nm in ("csi", "eta", "zet", "aj", "dtev"):
            assert np.array_equal(o.get(nm), z[nm]), nm
        return
    q = o.get("q")[:n1]
    assert np.array_equal(q, z["q"]), "q differs from the reference"
    assert np.array_equal(o.get("xnut")[:n1], z["xnut"])
    if "rh" in z.files:
        assert np.array_equal(o.get("rh")[:n1], z["rh"])
    if "conver" in z.files:
        check_conver(rows, z["conver"])


@pytest.mark.skipif(not po.ref_available(), reason="reference harness not built (needs /root/reference)")
@pytest.mark.parametrize("ng,des,bt", [(2, True, (1, 3, 1, 0, 2, 3)), (3, True, (6, 6, 1, 1, 6, 6)),
                                       (3, False, (0, 0, 0, 0, 0, 0)), (2, True, (4, 3, 1, 1, 2, 2)),
                                       (3, True, (3, 4, 4, 4, 4, 4)),
                                       # characteristic outlet (btype 5) on i = max, on both i faces,
                                       # and on j / k faces (nu~ extrapolation, no semi-explicit BC)
                                       (2, True, (3, 5, 1, 1, 2, 2)), (3, False, (5, 5, 1, 1, 2, 2)),
                                       (3, True, (4, 5, 5, 1, 2, 5))])
def test_oracle_matches_reference_live(ng, des, bt):
    c = cases.wavy(17, 17, 21, ng=ng, itm=3, des=des, btype=bt, iters=(1, 2, 3), seed=99, amp=0.06)
    o = po.Oracle(c.params).load_case(c)
    o.mg_driver()
    r = po.run_reference(c, mode=1, nsteps=1)
    q = o.get_state()[0]
    assert np.array_equal(q, r["q"])
    assert np.array_equal(o.get_state()[3], r["xnut"])


@pytest.mark.skipif(not po.ref_available(), reason="reference harness not built (needs /root/reference)")
@pytest.mark.parametrize("opts", [dict(mode=2, recalc=0, calc_pk=1), dict(mode=2, recalc=1, calc_pk=1),
                                  dict(mode=1, nsteps=2)])
def test_oracle_alm_matches_reference_live(opts):
    from desalm import alm
    c = cases.wavy(25, 17, 21, ng=2, itm=3, amp=0.03, des=True, iters=(1, 2, 3), seed=13)
    c.alm = alm.rotor(c, i_rot=11, center=(0.45, 0.55), radius=0.28, hub=0.05, nblades=2, nelems=6,
                      nangles=37, angle=1.1, tsr=5.0)
    o, _ = _run_oracle(c, opts)
    r = po.run_reference(c, **opts)
    assert np.array_equal(o.get_state()[0], r["q"])
    if opts["mode"] == 2:
        assert np.array_equal(o.get("rh")[:c.n1], r["rh"])


@pytest.mark.skipif(not po.ref_available(), reason="reference harness not built (needs /root/reference)")
@pytest.mark.parametrize("mode", [5, 6])
def test_oracle_bintp_matches_reference_live(mode):
    c = cases.wavy(15, 13, 11, ng=2, des=True, overset=70, seed=31)
    o, _ = _run_oracle(c, dict(mode=mode))
    r = po.run_reference(c, mode=mode)
    assert np.array_equal(o.get_state()[0], r["q"])
    assert np.array_equal(o.get_state()[3], r["xnut"])


@pytest.mark.skipif(not po.ref_available(), reason="reference harness not built (needs /root/reference)")
def test_oracle_overset_mg_driver_and_norms_live():
    c = cases.wavy(17, 17, 21, ng=3, itm=6, it_min=1, er_min=-0.08, des=True, iters=(1, 2, 3), seed=5,
                   overset=90, apply_mom=False)
    o, rows = _run_oracle(c, dict(mode=1, nsteps=2))
    r = po.run_reference(c, mode=1, nsteps=2)
    assert np.array_equal(o.get_state()[0], r["q"])
    assert np.array_equal(o.get_state()[3], r["xnut"])
    check_conver(rows, conver_table(r["conver"]))


@pytest.mark.skipif(not po.hooks_available(), reason="reference harness with hooks not built (needs /root/reference)")
@pytest.mark.parametrize("ng,bt,hooks", [(2, (3, 3, 1, 1, 2, 2), 3), (3, (0, 0, 1, 1, 2, 2), 3),
                                         (2, (3, 3, 1, 1, 2, 2), 1)])
def test_oracle_hooks_match_reference_live(ng, bt, hooks):
    """conservation and wf_eta (the reference's own Newton chain for u_tau)
    against the reference built with src/conservation.F90 and src/wf_mpi.F90.
    With only conservation on, the reference's wall functions are still
    linked in, so that case checks that the oracle's flags select them."""
    c = cases.wavy(17, 17, 21, ng=ng, itm=3, des=True, btype=bt, iters=(1, 2, 3), seed=5, amp=0.05)
    o = po.Oracle(c.params).load_case(c).set_hooks(3, wf_newton=True)
    o.mg_driver()
    r = po.run_reference(c, mode=1, nsteps=1, exe=po.HOOKS_EXE)
    assert np.array_equal(o.get_state()[0], r["q"])
    assert np.array_equal(o.get_state()[3], r["xnut"])
    if hooks != 3:
        o2 = po.Oracle(c.params).load_case(c).set_hooks(hooks, wf_newton=True)
        o2.mg_driver()
        assert not np.array_equal(o2.get_state()[0], r["q"])


@pytest.mark.skipif(not po.hooks_available(), reason="reference harness with hooks not built (needs /root/reference)")
def test_oracle_everything_together_matches_reference_live():
    """Actuator line, inflow, characteristic outlet, j walls and both hooks
    (the reference's Newton chain) in one run, against the reference built
    with conservation.F90 and wf_mpi.F90."""
    c = build("wavy", dict(ni=21, nj=17, nk=17, ng=2, itm=3, amp=0.02, des=True, iters=(1, 2, 3),
                           btype=(4, 5, 1, 1, 2, 2), alm=dict(i_rot=9, center=[0.5, 0.5], radius=0.3, hub=0.06)))
    o = po.Oracle(c.params).load_case(c).set_hooks(3, wf_newton=True)
    o.mg_driver()
    r = po.run_reference(c, mode=1, nsteps=1, exe=po.HOOKS_EXE)
    assert np.array_equal(o.get_state()[0], r["q"])
    assert np.array_equal(o.get_state()[3], r["xnut"])
