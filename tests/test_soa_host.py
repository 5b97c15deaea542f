"""The device layout builder (ccfd_amd/csrc/soa.cpp), built on the host and
checked against the preprocessed mesh with numpy: the Hilbert renumbering
is a permutation whose quadrants at the first two levels are contiguous and
whose consecutive elements are close; every slot's link names the
renumbered neighbour (plus the arena offset) and a partner slot that links
back, ghosts are -(g + offset + 1) and missing sides kNoSide; face lengths
carry the owner's sign; the ghost arrays point at their element and slot;
and each block's halo list is exactly the sorted set of face neighbours
outside the block. Every kernel indexes through these arrays. Host only."""
import ctypes as C
import os

import numpy as np
import pytest

from ccfd_amd import mesh as M
from tests import cases

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native",
                   "libsoa_host.so")
NO_SIDE = -2 ** 31
BLOCK = 256
INTS = ["perm", "inv", "link", "ghostBC", "ghostCell", "ghostK", "haloStart",
        "haloCells"]
DBLS = ["w", "gp", "fn", "flen", "bbv", "bbd", "areaq", "area", "sx", "sy",
        "bary", "ghostBary"]


def _lib():
    if not os.path.exists(LIB):
        pytest.skip("libsoa_host.so not built")
    L = C.CDLL(LIB)
    L.soa_build.restype = C.c_void_p
    L.soa_build.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]
    L.soa_free.argtypes = [C.c_void_p]
    L.soa_sizes.argtypes = [C.c_void_p, C.POINTER(C.c_long)]
    for f in ("soa_int_len", "soa_dbl_len"):
        getattr(L, f).restype = C.c_long
        getattr(L, f).argtypes = [C.c_void_p, C.c_int]
    L.soa_get_int.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
    L.soa_get_dbl.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
    return L


def build(mesh, S, eb=0, gb=0, bb=0):
    L = _lib()
    h = L.soa_build(C.cast(mesh.ptr, C.c_void_p), S, eb, gb, bb)
    sz = (C.c_long * 7)()
    L.soa_sizes(h, sz)
    o = dict(zip(["N", "Np", "S", "G", "nb", "maxHalo", "nHalo"], list(sz)))
    for k, name in enumerate(INTS):
        a = np.zeros(L.soa_int_len(h, k), np.int32)
        L.soa_get_int(h, k, a.ctypes.data_as(C.c_void_p))
        o[name] = a
    for k, name in enumerate(DBLS):
        a = np.zeros(L.soa_dbl_len(h, k))
        L.soa_get_dbl(h, k, a.ctypes.data_as(C.c_void_p))
        o[name] = a
    L.soa_free(h)
    return o


MESHES = {
    "delaunay_open": lambda: (cases.delaunay_mesh(34, seed=3, periodic=False), 3),
    "delaunay_periodic": lambda: (cases.delaunay_mesh(30, seed=6), 3),
    "mixed_tri_quad": lambda: (cases.mixed_mesh()[0], 4),
    "every_bc": lambda: (cases.mixed_bc_mesh(40, 23)[0], 4),
}


@pytest.mark.parametrize("offsets", [(0, 0, 0), (3 * BLOCK, 17, 5)])
@pytest.mark.parametrize("name", sorted(MESHES))
def test_layout_matches_the_mesh(name, offsets):
    mesh, S = MESHES[name]()
    eb, gb, bb = offsets
    o = build(mesh, S, eb, gb, bb)
    N, Np, G = mesh.n_elems, o["Np"], mesh.n_ghosts
    assert o["N"] == N and Np % BLOCK == 0 and N <= Np < N + BLOCK
    assert o["nb"] == Np // BLOCK and o["G"] == G
    perm, inv = o["perm"], o["inv"]
    assert np.array_equal(np.sort(perm), np.arange(N))
    assert np.array_equal(inv[perm], np.arange(N))

    # element arrays follow the renumbering
    assert np.array_equal(o["area"][:N], mesh.area[perm])
    assert np.array_equal(o["areaq"][:N], mesh.areaq[perm])
    assert np.array_equal(o["bary"][:N], mesh.bary[perm, 0])
    assert np.array_equal(o["bary"][Np:Np + N], mesh.bary[perm, 1])

    ss, nbr, part = mesh.slot_start, mesh.nbr, mesh.partner
    link = o["link"].reshape(S, Np)
    sl = S * Np
    for i in range(N):
        e = perm[i]
        s0, ns = ss[e], ss[e + 1] - ss[e]
        assert np.all(link[ns:, i] == NO_SIDE)
        for k in range(ns):
            s = s0 + k
            l = link[k, i]
            c = k * Np + i
            if nbr[s] >= 0:
                j = (l >> 2) - eb
                assert j == inv[nbr[s]]
                ps = l & 3
                assert part[s] == ss[nbr[s]] + ps
                assert link[ps, j] >> 2 == i + eb       # links back
                sign = 1.0 if mesh.owner[s] else -1.0
            else:
                assert l == nbr[s] - gb                  # -(g + gb + 1)
                sign = 1.0
            assert o["flen"][c] == sign * mesh.flen[s]
            assert o["w"][c] == mesh.w[s, 0] and o["w"][sl + c] == mesh.w[s, 1]
            assert o["gp"][c] == mesh.GP[s, 0] and o["fn"][sl + c] == mesh.fn[s, 1]

    # ghosts: element and slot of the boundary side, BC offset
    for g in range(G):
        s = mesh.ghost_slot[g]
        e = np.searchsorted(ss, s, side="right") - 1
        assert o["ghostCell"][g] == inv[e] + eb
        assert o["ghostK"][g] == s - ss[e]
        assert o["ghostBC"][g] == mesh.ghost_bc[g] + bb
        assert link[s - ss[e], inv[e]] == -(g + gb + 1)

    # halo lists, built again from the links
    hs, hc = o["haloStart"], o["haloCells"]
    assert hs[0] == 0 and hs[-1] == len(hc) == o["nHalo"]
    sizes = []
    for b in range(o["nb"]):
        lo, hi = b * BLOCK, min(N, (b + 1) * BLOCK)
        ln = link[:, lo:hi].ravel()
        j = (ln[ln >= 0] >> 2) - eb
        want = np.unique(j[(j < b * BLOCK) | (j >= (b + 1) * BLOCK)]) + eb
        assert np.array_equal(hc[hs[b]:hs[b + 1]], want), b
        sizes.append(len(want))
    assert o["maxHalo"] == max(sizes)


def test_hilbert_order_is_local():
    mesh = cases.delaunay_mesh(60, seed=8)
    o = build(mesh, 3)
    p = mesh.bary[o["perm"]]
    x0, y0 = mesh.bary.min(axis=0)
    span = (mesh.bary.max(axis=0) - [x0, y0]).max()
    # quadrants of the bounding square, two levels down, are visited as
    # contiguous runs
    for level in (1, 2):
        q = np.floor((p - [x0, y0]) / span * 2 ** level * (1 - 1e-12))
        cell = (q[:, 0] * 2 ** level + q[:, 1]).astype(int)
        runs = 1 + np.count_nonzero(np.diff(cell))
        assert runs == len(np.unique(cell)), level
    # consecutive elements are neighbours in space: the mean step is of the
    # order of the mesh spacing, far below that of the original order
    step = np.linalg.norm(np.diff(p, axis=0), axis=1).mean()
    h = 1.0 / 60
    assert step < 1.5 * h
    assert np.linalg.norm(np.diff(mesh.bary, axis=0), axis=1).mean() > 3 * step
