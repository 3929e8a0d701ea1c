"""desalm_gtsv_batch_noalloc, the allocation-free variant of
desalm_gtsv_batch, called through the C ABI: same results as the allocating
entry point on two calls in a row (the second smaller, so stale contents of
the shared device scratch would show), INFO of singular systems included,
and a clear error for batches beyond its fixed scratch."""
import numpy as np
import pytest

from desalm import lib as desalm_lib
from desalm.lib import DesalmError, gtsv_batch

MAX_ENTRIES = 262144  # DESALM_GTSV_NOALLOC_MAX of include/desalm.h

pytestmark = pytest.mark.gpu


def systems(rng, nsys, n, singular=()):
    d = rng.uniform(-1, 1, (nsys, n))
    dl = rng.uniform(-1, 1, (nsys, n))
    du = rng.uniform(-1, 1, (nsys, n))
    b = rng.standard_normal((nsys, n))
    for s in singular:
        # a zero first column: SGTSV stops with INFO = 1
        d[s, 0] = 0.0
        dl[s, 0] = 0.0
    return dl, d, du, b


def test_noalloc_matches_allocating_entry_twice(gpu):
    rng = np.random.default_rng(2027)
    for nsys, n, singular in ((1024, 200, (3, 700)), (96, 37, (5,))):
        dl, d, du, b = systems(rng, nsys, n, singular)
        x_ref, info_ref = gtsv_batch(dl, d, du, b)
        x, info = gtsv_batch(dl, d, du, b, noalloc=True)
        assert np.array_equal(info, info_ref)
        assert all(info[s] == 1 for s in singular)
        ok = info_ref == 0
        assert np.array_equal(x[ok], x_ref[ok])
        assert np.all(np.isfinite(x[ok]))


def test_noalloc_at_its_limit_and_beyond(gpu):
    rng = np.random.default_rng(7)
    n = 256
    dl, d, du, b = systems(rng, MAX_ENTRIES // n, n)
    x_ref, info_ref = gtsv_batch(dl, d, du, b)
    x, info = gtsv_batch(dl, d, du, b, noalloc=True)
    assert np.array_equal(info, info_ref)
    assert np.array_equal(x[info_ref == 0], x_ref[info_ref == 0])
    dl, d, du, b = systems(rng, MAX_ENTRIES // n + 1, n)
    with pytest.raises(DesalmError, match="exceeds the fixed scratch"):
        gtsv_batch(dl, d, du, b, noalloc=True)
    assert "DESALM_GTSV_NOALLOC_MAX" in desalm_lib.lib().desalm_last_error(None).decode()
