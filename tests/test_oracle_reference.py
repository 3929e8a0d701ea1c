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

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))


def _case(meta):
    kw = {k: tuple(v) if isinstance(v, list) else v for k, v in meta["kwargs"].items()}
    return getattr(cases, meta["kind"])(**kw)


def _run_oracle(case, opts):
    o = po.Oracle(case.params).load_case(case)
    mode = opts["mode"]
    if mode == 1:
        for _ in range(opts.get("nsteps", 1)):
            o.mg_driver()
    elif mode == 2:
        o.solver_daf(1, opts["recalc"], opts["calc_pk"])
    elif mode == 3:
        o.des_eddy()
    return o


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_oracle_matches_reference_golden(path):
    z = np.load(path, allow_pickle=False)
    meta = json.loads(str(z["meta"]))
    import importlib
    mk = importlib.import_module("golden.make_golden") if False else None  # noqa: F841
    case = _case(meta)
    # the fixture was produced from exactly these inputs
    import hashlib
    h = hashlib.sha256()
    for a in (case.x, case.y, case.z, case.wd, case.q, case.qn, case.qnm1, case.xnut):
        h.update(np.ascontiguousarray(a, dtype=np.float64).tobytes())
    assert h.hexdigest() == meta["inputs_sha256"], "case builder changed; regenerate the fixtures"
    o = _run_oracle(case, meta["options"])
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


@pytest.mark.skipif(not po.ref_available(), reason="reference harness not built (needs /root/reference)")
@pytest.mark.parametrize("ng,des,bt", [(2, True, (1, 3, 1, 0, 2, 3)), (3, True, (6, 6, 1, 1, 6, 6)),
                                       (3, False, (0, 0, 0, 0, 0, 0))])
def test_oracle_matches_reference_live(ng, des, bt):
    c = cases.wavy(17, 17, 21, ng=ng, itm=3, des=des, btype=bt, iters=(1, 2, 3), seed=99, amp=0.06)
    o = po.Oracle(c.params).load_case(c)
    o.mg_driver()
    r = po.run_reference(c, mode=1, nsteps=1)
    q = o.get_state()[0]
    assert np.array_equal(q, r["q"])
    assert np.array_equal(o.get_state()[3], r["xnut"])
