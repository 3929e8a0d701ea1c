"""The per-step outputs of the physical time loop (src/main.F90:71-75) from
the device state: desalm_real_residual against the reference's formula
(real_residual, src/main.F90:214-300) evaluated in numpy on the state read
back, desalm_probe against that state at the monitor points, and the files
append_step_outputs writes."""
import numpy as np
import pytest

from desalm import cases, files
from desalm.lib import DesalmError, Solver

pytestmark = pytest.mark.gpu


def test_real_residual_and_probes_match_the_state(gpu, tmp_path):
    c = cases.alm_box(41, 25, 25, ng=2, itm=5, nelems=6, nangles=73, iters=(1, 2))
    s = Solver(c.params).load_case(c)
    s.step()
    q, qn, _, xnut = s.get_state()
    ni, nj, nk = c.params.img, c.params.jmg, c.params.kmg
    idx = np.arange(c.n1)
    i, j, k = idx % ni, (idx // ni) % nj, idx // (ni * nj)
    inner = (i > 0) & (i < ni - 1) & (j > 0) & (j < nj - 1) & (k > 0) & (k < nk - 1)
    tnu = (ni - 2) * (nj - 2) * (nk - 2)
    want = np.log10(np.sqrt(((q[inner] - qn[inner]) ** 2).sum(axis=0) / tnu))
    got = s.real_residual()
    assert np.max(np.abs(got - want)) < 1e-12
    pts = np.array([[2, 2, 2], [ni // 2, nj // 2, nk // 2], [ni, nj, nk], [1, 1, 1]])
    pv = s.probe(pts)
    for p, (a, b, cc) in enumerate(pts):
        l = (a - 1) + ni * ((b - 1) + nj * (cc - 1))
        assert np.array_equal(pv[p, :5], q[l]) and pv[p, 5] == xnut[l]
    with pytest.raises(DesalmError, match="outside the fine level"):
        s.probe([[ni + 1, 1, 1]])
    files.append_step_outputs(str(tmp_path), 1, s, monitor_points=pts[:2], n_turbines=1)
    files.append_step_outputs(str(tmp_path), 2, s, monitor_points=pts[:2], n_turbines=1)
    assert len((tmp_path / "conver_real").read_text().splitlines()) == 2
    assert len((tmp_path / "history").read_text().splitlines()) == 4
    assert (tmp_path / "resultsCTCP01.dat").read_text().splitlines()[1].startswith("00002")
    s.close()
