"""ccfd_gpu_riemann_batch_noalloc: convectiveFlux on n state pairs staged
through the library's preallocated device array, with no allocation. Its
fluxes equal those of ccfd_gpu_riemann_batch (same kernel) bit for bit, on
two calls in a row, the second with fewer states (the array keeps the first
call's contents), in both precisions; sizes beyond the array are refused."""
import ctypes as C

import numpy as np
import pytest

from ccfd_amd import _native as N
from ccfd_amd import gpu as G

pytestmark = pytest.mark.gpu


def _states(n, seed):
    rng = np.random.default_rng(seed)
    s = np.empty((n, 4))
    s[:, 0] = rng.uniform(0.1, 3.0, n)
    s[:, 1] = rng.uniform(-2.0, 2.0, n)
    s[:, 2] = rng.uniform(-1.0, 1.0, n)
    s[:, 3] = rng.uniform(0.1, 3.0, n)
    return s


def _noalloc(code, prec, L, R):
    lib = G.lib()
    f = lib.ccfd_gpu_riemann_batch_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_int, C.c_int, C.c_double, C.c_long, N.c_double_p,
                  N.c_double_p, N.c_double_p]
    F = np.empty_like(L)
    G.check(f(code, prec, 1.4, len(L), N.dptr(L), N.dptr(R), N.dptr(F)))
    return F


@pytest.mark.parametrize("precision", [G.FP64, G.FP32_FLUX])
@pytest.mark.parametrize("code", [2, 3, 5, 9])
def test_noalloc_matches_the_allocating_entry_point(code, precision):
    for k, n in enumerate((12000, 700)):   # the second call is smaller
        L = _states(n, 10 * code + k)
        R = _states(n, 10 * code + k + 5)
        got = _noalloc(code, precision, L, R)
        want = G.riemann_batch(code, L, R, precision=precision)
        assert np.array_equal(got, want), (code, n)


def test_noalloc_refuses_more_states_than_its_array():
    n = 16384 + 1
    L = _states(n, 1)
    with pytest.raises(G.GpuError, match="exceed"):
        _noalloc(5, G.FP64, L, L)
    # and the array still serves a call that fits
    L = _states(16384, 2)
    R = _states(16384, 3)
    assert np.array_equal(_noalloc(5, G.FP64, L, R), G.riemann_batch(5, L, R))
