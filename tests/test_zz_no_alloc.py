"""cf_eqwm_batch_noalloc (staging in a fixed device array, no device
allocation) against cf_eqwm_batch (per-call device allocations) through the
C-ABI: the same kernel on the same inputs, so the results are bit-identical.
Two calls in a row, the second smaller, so stale contents of the shared
array from the first call cannot leak into the second."""
import numpy as np
import pytest

from channel_flow import _lib

pytestmark = pytest.mark.gpu


def _inputs(n, seed):
    rng = np.random.default_rng(seed)
    u = rng.uniform(0.2, 2.0, n)
    nu = 10 ** rng.uniform(-5, -3, n)
    h = rng.uniform(0.05, 0.2, n)
    tin = rng.uniform(1e-4, 1e-2, n)
    return [np.ascontiguousarray(a) for a in (u, nu, h, tin)]


def _call(fn, cols, ncv, stretch, tol, max_iter):
    n = cols[0].size
    tau = np.full(n, np.nan)
    it = np.full(n, -1, dtype=np.int32)
    rc = fn(n, *(_lib.dptr(c) for c in cols), ncv, stretch, tol, max_iter, _lib.dptr(tau), _lib.iptr(it))
    return rc, tau, it


def test_noalloc_matches_allocating_entry_point():
    lib = _lib.load()
    for n, seed, ncv, tol in ((40000, 11, 30, 1e-8), (777, 12, 100, 1e-7)):
        print("eqwm noalloc case n=%d seed=%d ncv=%d" % (n, seed, ncv), flush=True)
        cols = _inputs(n, seed)
        rc_a, tau_a, it_a = _call(lib.cf_eqwm_batch, cols, ncv, 1.1, tol, 100)
        _lib.check(rc_a, "cf_eqwm_batch")
        rc_b, tau_b, it_b = _call(lib.cf_eqwm_batch_noalloc, cols, ncv, 1.1, tol, 100)
        _lib.check(rc_b, "cf_eqwm_batch_noalloc")
        np.testing.assert_array_equal(tau_b, tau_a)
        np.testing.assert_array_equal(it_b, it_a)
        assert np.all(np.isfinite(tau_b)) and np.all(it_b >= 0)


def test_noalloc_rejects_oversized_batch():
    """n above CF_EQWM_NOALLOC_MAX (524288) is refused on the host, before
    any array is read."""
    lib = _lib.load()
    big = 524288 + 1
    one = np.ones(1)
    tau = np.zeros(1)
    it = np.zeros(1, dtype=np.int32)
    rc = lib.cf_eqwm_batch_noalloc(big, _lib.dptr(one), _lib.dptr(one), _lib.dptr(one), _lib.dptr(one), 30, 1.1,
                                   1e-8, 100, _lib.dptr(tau), _lib.iptr(it))
    assert rc != 0
    assert b"exceeds the maximum" in lib.cf_last_error()
