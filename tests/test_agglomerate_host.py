"""Element agglomeration for the multigrid preconditioner of the implicit
solver (ccfd_amd/csrc/agglomerate.cpp), built on the host.

The coarse blocks are sums of the blocks of the level above, as listed by
the agglomeration; the test assembles a random block matrix, sums it as
listed and compares with the Galerkin product P^T A P built independently
(scipy.sparse, piecewise-constant P from the group map), level by level.
A dropped, doubled or misplaced contribution fails it."""
import ctypes as C
import math
import os

import numpy as np
import pytest
import scipy.sparse as sp

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native",
                   "libagglomerate_host.so")


def _lib():
    if not os.path.exists(LIB):
        pytest.skip("libagglomerate_host.so not built")
    L = C.CDLL(LIB)
    ip, dp = C.POINTER(C.c_int), C.POINTER(C.c_double)
    L.agg_build.argtypes = [C.c_int, C.c_int, C.c_int, ip, dp, C.c_int, C.c_int]
    L.agg_build.restype = C.c_void_p
    L.agg_free.argtypes = [C.c_void_p]
    L.agg_levels.argtypes = [C.c_void_p]
    L.agg_sizes.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_long)]
    L.agg_get.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p]
    return L


NAMES = ["aggOf", "rowStart", "col", "dSrcStart", "dSrc", "oSrcStart",
         "oSrc", "colorStart", "colorCells"]


def build(link, w, min_cells=16, max_levels=10):
    L = _lib()
    S, n = link.shape
    link = np.ascontiguousarray(link, dtype=np.int32)
    w = np.ascontiguousarray(w, dtype=np.float64)
    h = L.agg_build(n, n, S, link.ctypes.data_as(C.POINTER(C.c_int)),
                    w.ctypes.data_as(C.POINTER(C.c_double)), min_cells,
                    max_levels)
    levels = []
    for l in range(L.agg_levels(h)):
        sz = (C.c_long * 9)()
        L.agg_sizes(h, l, sz)
        lev = {}
        for k, name in enumerate(NAMES):
            a = np.zeros(sz[k], dtype=np.int64 if 3 <= k <= 6 else np.int32)
            if sz[k]:
                L.agg_get(h, l, k, a.ctypes.data_as(C.c_void_p))
            lev[name] = a
        lev["n"] = len(lev["rowStart"]) - 1
        levels.append(lev)
    L.agg_free(h)
    return levels


def ring_grid(ni, nj, h0, growth, radius0=0.5, perm_seed=None):
    """ni x nj quads around a circle, periodic in i; radial spacing h0 *
    growth**j; slots 0: j-1 (wall side), 1: i+1, 2: j+1, 3: i-1. Returns
    link [4][N] (neighbour << 2 | partner slot, -1 at the wall and the far
    side), weights [4][N] (face length / centre distance) and the element
    index of (i, j)."""
    n = ni * nj
    idx = np.arange(n).reshape(nj, ni)
    if perm_seed is not None:
        idx = np.random.default_rng(perm_seed).permutation(n).reshape(nj, ni)
    hn = h0 * growth ** np.arange(nj)
    r = radius0 + np.concatenate([[0.0], np.cumsum(hn)])
    ht = 2 * math.pi * 0.5 * (r[:-1] + r[1:]) / ni
    link = -np.ones((4, n), dtype=np.int32)
    w = np.zeros((4, n))
    partner = [2, 3, 0, 1]
    for j in range(nj):
        for i in range(ni):
            e = idx[j, i]
            nb = [(i, j - 1), ((i + 1) % ni, j), (i, j + 1), ((i - 1) % ni, j)]
            for s, (a, b) in enumerate(nb):
                if 0 <= b < nj:
                    link[s, e] = (idx[b, a] << 2) | partner[s]
                    w[s, e] = (ht[j] / (0.5 * (hn[j] + hn[b])) if s in (0, 2)
                               else hn[j] / ht[j])
    return link, w, idx


def block_matrix(D, O, rows, cols, n):
    """sparse 4n x 4n matrix from diagonal blocks D [16][n] and off-diagonal
    blocks O [16][m] at (rows, cols)"""
    r, c, v = [], [], []
    for a in range(4):
        for b in range(4):
            r += [4 * np.arange(n) + a, 4 * rows + a]
            c += [4 * np.arange(n) + b, 4 * cols + b]
            v += [D[4 * a + b], O[4 * a + b]]
    return sp.csr_matrix((np.concatenate(v), (np.concatenate(r),
                                              np.concatenate(c))),
                         shape=(4 * n, 4 * n))


def sum_blocks(lev, Dp, Op):
    """coarse blocks summed as the agglomeration lists them"""
    def gather(start, src, m):
        out = np.zeros((16, m))
        for k in range(m):
            for code in src[start[k]:start[k + 1]]:
                idx = code >> 1
                out[:, k] += Dp[:, idx] if code & 1 else Op[:, idx]
        return out
    return (gather(lev["dSrcStart"], lev["dSrc"], lev["n"]),
            gather(lev["oSrcStart"], lev["oSrc"], len(lev["col"])))


def check_levels(link, levels, rng):
    S, n = link.shape
    # level-0 random blocks: diagonal [16][n], side blocks stored as s*n+i
    D = rng.normal(size=(16, n))
    Os = rng.normal(size=(16, S * n))
    e = np.nonzero((link >= 0) & ((link >> 2) != np.arange(n)[None, :]))
    rows0, cols0 = e[1], link[e] >> 2
    A = block_matrix(D, Os[:, e[0] * n + e[1]], rows0, cols0, n)
    Dp, Op, nprev = D, Os, n
    for lev in levels:
        m = lev["n"]
        agg = lev["aggOf"]
        assert len(agg) == nprev and agg.min() == 0 and agg.max() == m - 1
        sizes = np.bincount(agg, minlength=m)
        assert sizes.min() >= 1 and sizes.max() <= 4
        assert m <= nprev - nprev // 8
        rs, col = lev["rowStart"], lev["col"]
        rows = np.repeat(np.arange(m), np.diff(rs))
        assert np.all(col != rows)
        Dc, Oc = sum_blocks(lev, Dp, Op)
        # Galerkin product with piecewise-constant prolongation
        P1 = sp.csr_matrix((np.ones(nprev), (np.arange(nprev), agg)),
                           shape=(nprev, m))
        P = sp.kron(P1, sp.identity(4), format="csr")
        Ac = (P.T @ A @ P).toarray()
        Bc = block_matrix(Dc, Oc, rows, col, m).toarray()
        assert np.allclose(Bc, Ac, rtol=1e-12, atol=1e-12)
        # colouring: no two neighbours share a colour
        color = np.empty(m, np.int64)
        cs, cc = lev["colorStart"], lev["colorCells"]
        for k in range(len(cs) - 1):
            color[cc[cs[k]:cs[k + 1]]] = k
        assert np.all(color[rows] != color[col])
        A, Dp, Op, nprev = sp.csr_matrix(Ac), Dc, Oc, m
    return nprev


def test_galerkin_blocks_on_a_stretched_ring():
    link, w, idx = ring_grid(32, 24, 1e-3, 1.2, perm_seed=5)
    levels = build(link, w)
    assert len(levels) >= 3
    coarsest = check_levels(link, levels, np.random.default_rng(1))
    assert coarsest <= 16 or len(levels) == 10
    # the first groups pair elements across the stretched direction: near
    # the wall, groups span rows (wall-normal), hardly ever columns
    agg = levels[0]["aggOf"]
    jj, ii = np.nonzero(idx >= 0)
    rows_of = {}
    for j, i in zip(jj, ii):
        rows_of.setdefault(agg[idx[j, i]], []).append((j, i))
    wall = [g for g in rows_of.values() if min(j for j, _ in g) <= 2]
    normal = sum(len({j for j, _ in g}) > len({i for _, i in g}) for g in wall)
    print("wall groups spanning rows: %d of %d" % (normal, len(wall)))
    assert normal >= 0.9 * len(wall)


def test_galerkin_blocks_on_an_irregular_graph():
    # a ring grid with random, symmetric face weights
    rng = np.random.default_rng(7)
    link, _, _ = ring_grid(20, 15, 1.0, 1.0, perm_seed=2)
    w = np.zeros(link.shape)
    for s in range(4):
        for e in np.nonzero(link[s] >= 0)[0]:
            j, ps = link[s, e] >> 2, link[s, e] & 3
            if w[s, e] == 0:
                w[s, e] = w[ps, j] = math.exp(rng.uniform(-3, 3))
    levels = build(link, w, min_cells=4)
    check_levels(link, levels, rng)


def test_small_mesh_has_no_levels():
    link, w, _ = ring_grid(4, 3, 1.0, 1.0)
    assert build(link, w, min_cells=16) == []
