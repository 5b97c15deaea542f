"""Host parts of the analysis consumers against the oracle, on the CPU:
the record-point search of the host library (ccfd_mesh_locate) and the
host exact functions used for initial conditions (ccfd_exact_func), both
bit for bit; the latter are also checked against the reference's compiled
exactFunction.c where the harness is built."""
import ctypes as C

import numpy as np
import pytest

from ccfd_amd import gpu as G
from ccfd_amd import mesh as M
from ccfd_amd import _native as N
from oracle import bindings as B


def _locate(mesh, x):
    L = N.host()
    f = L.ccfd_mesh_locate
    f.restype = C.c_long
    f.argtypes = [C.c_void_p, N.c_double_p]
    x = np.ascontiguousarray(x, dtype=np.float64)
    return f(C.cast(mesh.ptr, C.c_void_p), N.dptr(x))


@pytest.mark.parametrize("kind", ["delaunay", "cartesian", "ogrid"])
def test_record_point_search_matches_the_oracle(kind):
    if kind == "delaunay":
        raw, bcs = M.jittered_delaunay(18, seed=5, periodic=False)
        mesh = M.Mesh(raw, bcs)
    elif kind == "cartesian":
        raw = M.cartesian(10, 7, (0.0, 0.0), (2.0, 1.0), (101, 401, 101, 401))
        mesh = M.Mesh(raw, [M.BC(101), M.BC(401)])
    else:
        raw = M.naca_ogrid(ni=48, nj=10, wall_spacing=5e-3, radius=4.0)
        mesh = M.Mesh(raw, [M.BC(101), M.BC(501, rho=1.0, mach=0.5,
                                                 alpha=0.0, pressure=1.0)])
    rng = np.random.default_rng(2)
    lo, hi = raw.xy.min(0), raw.xy.max(0)
    pts = lo + (hi - lo) * rng.random((60, 2))
    pts = np.vstack([pts, raw.xy[:5] + 1e-9, hi + 1.0])
    # points outside the domain included: a mesh whose boundary has
    # horizontal edges may accept them (the reference's 0 / 0), and both
    # implementations must agree on that too
    for x in pts:
        assert _locate(mesh, x) == B.locate(mesh, x)


def _exact_host(f, prm, mesh, x, t):
    L = G.lib()
    fn = L.ccfd_exact_func
    fn.restype = C.c_int
    fn.argtypes = [C.c_int, C.POINTER(G.Params), C.c_void_p, N.c_double_p,
                   C.c_double, N.c_double_p]
    out = np.zeros(4)
    x = np.ascontiguousarray(x, dtype=np.float64)
    rc = fn(f, C.byref(prm), C.cast(mesh.ptr, C.c_void_p), N.dptr(x), t,
            N.dptr(out))
    return rc, out


def test_host_exact_functions_match_the_oracle():
    raw = M.cartesian(16, 10, (-0.3, 0.1), (1.7, 1.3), (101, 401, 101, 401))
    mesh = M.Mesh(raw, [M.BC(101), M.BC(401)])
    left, right = (1.0, 0.75, 0.0, 1.0), (0.125, 0.0, 0.0, 0.1)
    prm = G.params(left=left, right=right, x0=0.4)
    cfg = B.make_cfg(left=left, right=right, x0=0.4,
                     dxref=np.sqrt(1.0 / (mesh.total_area_q * mesh.n_elems)))
    rng = np.random.default_rng(4)
    for f in range(1, 7):
        for t in (0.0, 0.031, 0.2):
            for x in mesh.bary[rng.choice(mesh.n_elems, 12)]:
                rc, got = _exact_host(f, prm, mesh, x, t)
                assert rc == 0
                want = B.exact_func(cfg, mesh, f, x, t)
                assert np.array_equal(got, want), (f, t, x, got, want)
    assert _exact_host(7, prm, mesh, mesh.bary[0], 0.0)[0] != 0
