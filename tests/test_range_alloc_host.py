"""The batched arena's first-fit range allocator (element blocks, ghosts,
halo lists and boundary-condition entries of all instances;
ccfd_amd/csrc/range_alloc.hpp), built on the host and driven through random
alloc / release sequences against a bitmap model: the same first-fit start,
the same answer to "does n fit", the same high-water mark and use count,
and free ranges that are always coalesced. Overlapping instances in HBM
would follow from any mismatch. Host only."""
import ctypes as C
import os

import numpy as np
import pytest

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native",
                   "librange_alloc_host.so")


def _lib():
    if not os.path.exists(LIB):
        pytest.skip("librange_alloc_host.so not built")
    L = C.CDLL(LIB)
    L.ra_new.restype = C.c_void_p
    L.ra_new.argtypes = [C.c_int]
    for f in ("ra_delete", "ra_high", "ra_used", "ra_ranges"):
        getattr(L, f).argtypes = [C.c_void_p]
    L.ra_alloc.argtypes = [C.c_void_p, C.c_int]
    L.ra_can.argtypes = [C.c_void_p, C.c_int]
    L.ra_release.argtypes = [C.c_void_p, C.c_int, C.c_int]
    return L


class Model:
    def __init__(self, cap):
        self.busy = np.zeros(cap, bool)

    def first_fit(self, n):
        if n == 0:
            return 0
        free = ~self.busy
        run = 0
        for i, f in enumerate(free):
            run = run + 1 if f else 0
            if run == n:
                return i - n + 1
        return -1

    def runs(self):
        """number of maximal free runs"""
        f = (~self.busy).astype(int)
        return int(np.sum(np.diff(np.concatenate([[0], f])) == 1))

    def high(self):
        idx = np.nonzero(self.busy)[0]
        return int(idx[-1] + 1) if len(idx) else 0


@pytest.mark.parametrize("seed", range(6))
def test_matches_a_bitmap_model(seed):
    L = _lib()
    rng = np.random.default_rng(seed)
    cap = int(rng.integers(40, 400))
    h = L.ra_new(cap)
    m = Model(cap)
    live = []
    try:
        for step in range(1500):
            if live and rng.random() < 0.45:
                s, n = live.pop(int(rng.integers(len(live))))
                L.ra_release(h, s, n)
                m.busy[s:s + n] = False
            else:
                n = int(rng.integers(0, cap // 4 + 2))
                want = m.first_fit(n)
                assert bool(L.ra_can(h, n)) == (want >= 0), (step, n)
                got = L.ra_alloc(h, n)
                assert got == want, (step, n, got, want)
                if got >= 0 and n > 0:
                    assert not m.busy[got:got + n].any()
                    m.busy[got:got + n] = True
                    live.append((got, n))
            assert L.ra_used(h) == int(m.busy.sum()), step
            assert L.ra_high(h) == m.high(), step
            assert L.ra_ranges(h) == m.runs(), step
        for s, n in live:                   # all back: one range again
            L.ra_release(h, s, n)
        assert L.ra_used(h) == 0 and L.ra_high(h) == 0
        assert L.ra_ranges(h) == 1 and L.ra_alloc(h, cap) == 0
    finally:
        L.ra_delete(h)


def test_full_and_empty_edges():
    L = _lib()
    h = L.ra_new(10)
    try:
        assert L.ra_alloc(h, 10) == 0 and L.ra_high(h) == 10
        assert L.ra_ranges(h) == 0 and not L.ra_can(h, 1)
        assert L.ra_alloc(h, 1) == -1 and L.ra_alloc(h, 0) == 0
        L.ra_release(h, 3, 2)               # a hole in the middle
        assert L.ra_high(h) == 10 and L.ra_alloc(h, 3) == -1
        assert L.ra_alloc(h, 2) == 3
    finally:
        L.ra_delete(h)
    h = L.ra_new(0)
    try:
        assert L.ra_alloc(h, 1) == -1 and L.ra_high(h) == 0
    finally:
        L.ra_delete(h)
