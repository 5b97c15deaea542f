"""The oracle's analysis consumers and source term (src/analyze.c,
src/source.c), pinned on the CPU.

* source term: bit for bit against the reference's compiled source.c and
  fvTimeDerivative (the harness), Euler and Navier-Stokes, triangles and
  quadrilaterals;
* quadrature points: rebuilt from the element nodes (side midpoints, and
  the midpoint of nodes 0 and 2 of a quadrilateral), weights summing to
  the area;
* calcErrors: a uniform offset gives the offset in every norm; at 2nd order
  the norms equal a brute-force evaluation from the nodes, the (pinned)
  limited gradients and the closed-form exact function;
* calcCoef / initWing: on an airfoil O-grid whose wall cells carry a linear
  pressure field, the wall integral equals minus the body area times the
  gradient (the midpoint rule is exact for linear functions); lists are
  sorted by x, split by the normal's sign, and cover the wall once;
* initRecordPoints: the element found contains the point (Delaunay mesh),
  and on a Cartesian mesh horizontal edges never reject a point, as in the
  reference (its normalisation divides by zero for them).
"""
import os

import numpy as np
import pytest

from ccfd_amd import mesh as M
from oracle import bindings as B
from tests import cases


def _tri_mesh(n=14, seed=3):
    raw, bcs = M.jittered_delaunay(n, seed=seed, periodic=False, bc_code=101)
    return M.Mesh(raw, bcs)


def _quad_mesh(ni=9, nj=6):
    raw = M.cartesian(ni, nj, (0.0, 0.0), (1.0, 0.8), (101, 401, 101, 401))
    return M.Mesh(raw, [M.BC(101), M.BC(401)])


@pytest.mark.parametrize("kind", ["tri", "quad"])
def test_quadrature_points_from_the_nodes(kind):
    mesh = _tri_mesh() if kind == "tri" else _quad_mesh()
    xy = mesh.raw.xy
    for e in range(mesh.n_elems):
        x, w = B.quadrature(mesh, e)
        nodes = xy[mesh.elem_node[e][:mesh.elem_type[e]]]
        k = len(nodes)
        mids = 0.5 * (nodes + np.roll(nodes, -1, axis=0))
        # the side points come in the element's side order (the reference
        # prepends each new side), so compare them as sets
        def rows(a):
            return a[np.lexsort((a[:, 1], a[:, 0]))]
        if k == 3:
            assert len(x) == 3
            np.testing.assert_allclose(rows(x), rows(mids), rtol=0, atol=1e-14)
        else:
            assert len(x) == 5
            np.testing.assert_allclose(rows(x[:4]), rows(mids), rtol=0,
                                       atol=1e-14)
            np.testing.assert_allclose(x[4], 0.5 * (nodes[0] + nodes[2]),
                                       rtol=0, atol=1e-15)
        assert abs(w.sum() - mesh.area[e]) <= 1e-15 * mesh.area[e] * 4


def _ref_source_case(mesh, bcs_items, ns, order):
    ini = "/tmp/ccfd_src_%d.ini" % os.getpid()
    head = [("mu", 0.01), ("Pr", 0.72)] if ns else []
    B.write_ini(ini, head + [("fluxFunction", 2), ("calcSource", "T"),
                             ("sourceFunction", 1)] + bcs_items +
                [("spatialOrder", order), ("limiter", 2), ("venk_k", 5.0)])
    return ini


@pytest.mark.parametrize("kind,ns,order", [("tri", 0, 2), ("quad", 0, 1),
                                           ("quad", 1, 2)])
def test_source_term_matches_reference(kind, ns, order):
    """fvTimeDerivative with calcSource (sourceFunction 1) bit for bit
    against the reference's compiled source.c + finiteVolume.c"""
    if not B.ref_available(True):
        pytest.skip("reference harness not built")
    mesh = _tri_mesh() if kind == "tri" else _quad_mesh()
    items = [("nBC", len(mesh.bcs))] + [("BCtype", b.code) for b in mesh.bcs]
    ini = _ref_source_case(mesh, items, ns, order)
    prim = cases.random_state(mesh.n_elems, seed=5, mach=0.4)
    try:
        ref = B.Ref(ns=bool(ns))
        ref.setup(ini, mesh)
        ref.set_prim(prim)
        r = ref.fvtd(0.37)
    finally:
        os.unlink(ini)
    cfg = B.make_cfg(flux=2, order=order, limiter=2, venk_k=5.0,
                     navierstokes=ns, mu=0.01 if ns else 0.0, source=1)
    orc = B.Oracle(mesh, cfg)
    orc.set_prim(prim)
    ut = orc.fvtd(0.37)
    assert np.array_equal(ut, r["u_t"])
    # and the source changes dU/dt (the term is not silently zero)
    orc0 = B.Oracle(mesh, B.make_cfg(flux=2, order=order, limiter=2,
                                     venk_k=5.0, navierstokes=ns,
                                     mu=0.01 if ns else 0.0))
    orc0.set_prim(prim)
    assert np.abs(orc0.fvtd(0.37) - ut).max() > 1e-3


def test_errors_of_a_uniform_offset():
    """exact function 2 (pressure pulse: rho = 1, v = 0) against a state
    1 + d: every norm of rho, vx, vy is d at first order"""
    mesh = _quad_mesh()
    d = 0.03125
    prim = np.empty((mesh.n_elems, 4))
    prim[:] = (1.0 + d, d, -d, 1.0)
    orc = B.Oracle(mesh, B.make_cfg(order=1))
    orc.set_prim(prim)
    L1, L2, Li = orc.calc_errors(2, 0.0)
    for v in range(3):
        assert abs(L1[v] - d) < 1e-15 and abs(L2[v] - d) < 1e-15
        assert Li[v] == d


def _exact3(x, t):
    """exact function 3 (src/exactFunction.c:79-90) in closed form:
    rho = 2 + 0.1 sin(pi (x + y) - 2 pi t), v = (1, 1), p from E = rho^2"""
    r = 2.0 + 0.1 * np.sin(np.pi * (x[:, 0] + x[:, 1]) - 2 * np.pi * t)
    p = 0.4 * (r * r - 0.5 * r * 2.0)
    return np.stack([r, np.ones_like(r), np.ones_like(r), p], 1)


@pytest.mark.parametrize("kind", ["tri", "quad"])
def test_errors_against_brute_force(kind):
    mesh = _tri_mesh() if kind == "tri" else _quad_mesh()
    prim = cases.random_state(mesh.n_elems, seed=9, mach=0.3)
    prim[:, 0] += 1.0
    cfg = B.make_cfg(order=2, limiter=2, venk_k=3.0)
    orc = B.Oracle(mesh, cfg)
    orc.set_prim(prim)
    t = 0.21
    g = orc.fvtd(t, full=True)    # the pinned limited gradients at t
    L1, L2, Li = orc.calc_errors(3, t)
    xy = mesh.raw.xy
    acc1, acc2, accI = np.zeros(4), np.zeros(4), np.zeros(4)
    for e in range(mesh.n_elems):
        nodes = xy[mesh.elem_node[e][:mesh.elem_type[e]]]
        mids = 0.5 * (nodes + np.roll(nodes, -1, axis=0))
        if len(nodes) == 3:
            pts, w = mids, np.full(3, mesh.area[e] / 3)
        else:
            pts = np.vstack([mids, 0.5 * (nodes[0] + nodes[2])])
            w = np.array([1, 1, 1, 1, 2]) * mesh.area[e] / 6
        dx = pts - mesh.bary[e]
        rec = prim[e] + dx[:, :1] * g["u_x"][e] + dx[:, 1:] * g["u_y"][e]
        err = np.abs(_exact3(pts, t) - rec)
        acc1 += (err * w[:, None]).sum(0)
        acc2 += (err * err * w[:, None]).sum(0)
        accI = np.maximum(accI, err.max(0))
    A = mesh.area.sum()
    np.testing.assert_allclose(L1, acc1 / A, rtol=1e-12)
    np.testing.assert_allclose(L2, np.sqrt(acc2 / A), rtol=1e-12)
    np.testing.assert_allclose(Li, accI, rtol=1e-12)


def _wing_case(alpha=3.0):
    raw = M.naca_ogrid(ni=64, nj=12, wall_spacing=2e-3, radius=6.0)
    mesh = M.Mesh(raw, [M.BC(101), M.BC(501, rho=1.0, mach=0.5, alpha=alpha,
                                          pressure=1.0)])
    return raw, mesh


def test_wing_lists_cover_the_wall_sorted_and_split():
    raw, mesh = _wing_case()
    orc = B.Oracle(mesh, B.make_cfg(order=1))
    slots, nP, nS = orc.wing_lists(101)
    walls = mesh.ghost_slot[mesh.ghost_bc == 0]
    assert sorted(slots.tolist()) == sorted(walls.tolist())
    assert nP + nS == len(walls) and nP > 0 and nS > 0
    se = mesh.slot_elem()
    x = mesh.GP[slots, 0] + mesh.bary[se[slots], 0]
    assert np.all(np.diff(x[:nP]) >= 0) and np.all(np.diff(x[nP:]) >= 0)
    # suction list: ghost normal n_y <= 0, i.e. interior n_y >= 0
    assert np.all(mesh.n[slots[nP:], 1] >= 0)
    assert np.all(mesh.n[slots[:nP], 1] < 0)
    assert orc.wing_lists(999)[0].size == 0


def test_wing_coefficients_of_a_linear_pressure_field():
    """p = a + b y + c x on the wall midpoints: sum n p len over the wall
    (n out of the fluid, into the body) is -(c, b) times the body area"""
    alpha = 3.0
    raw, mesh = _wing_case(alpha)
    se = mesh.slot_elem()
    walls = mesh.ghost_slot[mesh.ghost_bc == 0]
    we = se[walls]
    assert len(np.unique(we)) == len(we)       # one wall side per cell
    a, b, c = 1.3, 0.25, -0.4
    mid = mesh.GP[walls] + mesh.bary[we]
    prim = np.empty((mesh.n_elems, 4))
    prim[:] = (1.0, 0.5, 0.0, 1.0)
    prim[we, 3] = a + b * mid[:, 1] + c * mid[:, 0]
    orc = B.Oracle(mesh, B.make_cfg(order=1))
    orc.set_prim(prim)
    orc.fvtd(0.0)
    slots, nP, nS = orc.wing_lists(101)
    gam = 1.4
    rho, p = 1.0, 1.0
    v = 0.5 * np.sqrt(gam * p / rho)
    ref = (rho, v * np.cos(np.radians(alpha)), v * np.sin(np.radians(alpha)), p)
    L = 0.75
    cl, cd, cp, xyphi = orc.wing_coef(slots, nP, nS, L, ref, alpha)
    # body area by the shoelace formula over the wall polygon (nodes)
    surf = raw.xy[:64]
    area = 0.5 * abs(np.dot(surf[:, 0], np.roll(surf[:, 1], -1)) -
                     np.dot(surf[:, 1], np.roll(surf[:, 0], -1)))
    q = 0.5 * rho * v * v
    cl0, cd0 = -b * area / (q * L), -c * area / (q * L)
    ar = np.radians(alpha)
    assert abs(cl - (cl0 * np.cos(ar) - cd0 * np.sin(ar))) < 1e-12
    assert abs(cd - (cd0 * np.cos(ar) + cl0 * np.sin(ar))) < 1e-12
    pw = a + b * xyphi[:, 1] + c * xyphi[:, 0]
    np.testing.assert_allclose(cp, (pw - p) / q, rtol=0, atol=1e-13)
    # numpy's arctan2 and the C library's atan2 may differ by an ulp
    np.testing.assert_allclose(xyphi[:, 2], np.arctan2(xyphi[:, 1],
                                                       xyphi[:, 0]),
                               rtol=4.5e-16, atol=0)


def _contains(nodes, x):
    k = len(nodes)
    for i in range(k):
        a, b = nodes[i], nodes[(i + 1) % k]
        if (b[0] - a[0]) * (x[1] - a[1]) - (b[1] - a[1]) * (x[0] - a[0]) < -1e-14:
            return False
    return True


def test_record_point_lies_in_its_element():
    mesh = _tri_mesh(n=20, seed=4)
    xy = mesh.raw.xy
    rng = np.random.default_rng(1)
    lo, hi = xy.min(0), xy.max(0)
    for x in lo + (hi - lo) * (0.1 + 0.8 * rng.random((40, 2))):
        e = B.locate(mesh, x)
        assert e >= 0
        nodes = xy[mesh.elem_node[e][:mesh.elem_type[e]]]
        # counter-clockwise or clockwise: the point is inside either way
        assert _contains(nodes, x) or _contains(nodes[::-1], x)


def test_record_point_on_cartesian_mesh_keeps_reference_behaviour():
    """horizontal edges give 0/0 in the reference's normalisation and never
    reject: the first element (in element order) whose vertical edges
    bracket x is taken, whatever its y range"""
    mesh = _quad_mesh(ni=8, nj=5)
    xy = mesh.raw.xy
    x = np.array([0.53, 0.71])
    e = B.locate(mesh, x)
    nodes = xy[mesh.elem_node[e][:4]]
    assert nodes[:, 0].min() <= x[0] <= nodes[:, 0].max()
    first = [k for k in range(mesh.n_elems)
             if xy[mesh.elem_node[k][:4], 0].min() <= x[0] <=
             xy[mesh.elem_node[k][:4], 0].max()][0]
    assert e == first
    assert not (nodes[:, 1].min() <= x[1] <= nodes[:, 1].max())


def test_exact_functions_match_reference():
    """oracle_exact_func against the reference's compiled exactFunction.c
    (the harness's globals: bounding box, dxRef, Riemann states), all six
    functions, bit for bit"""
    import ctypes as C
    if not B.ref_available():
        pytest.skip("reference harness not built")
    mesh = _quad_mesh(ni=11, nj=7)
    left, right = (1.0, 0.75, 0.0, 1.0), (0.125, 0.0, 0.0, 0.1)
    ini = "/tmp/ccfd_ex_%d.ini" % os.getpid()
    B.write_ini(ini, [("fluxFunction", 5), ("nBC", 2), ("BCtype", 101),
                      ("BCtype", 401), ("spatialOrder", 1)])
    try:
        ref = B.Ref()
        ref.setup(ini, mesh, left=left, right=right, x0=0.4)
    finally:
        os.unlink(ini)
    f = ref.L.exactFunc
    f.restype = None
    f.argtypes = [C.c_int, B.dp, C.c_double, B.dp]
    cfg = B.make_cfg(left=left, right=right, x0=0.4,
                     dxref=np.sqrt(1.0 / (mesh.total_area_q * mesh.n_elems)))
    rng = np.random.default_rng(3)
    for fn in range(1, 7):
        for t in (0.0, 0.043, 0.31):
            for x in mesh.bary[rng.choice(mesh.n_elems, 10)]:
                x = np.ascontiguousarray(x)
                want = np.zeros(4)
                f(fn, B._d(x), t, B._d(want))
                got = B.exact_func(cfg, mesh, fn, x, t)
                assert np.array_equal(got, want), (fn, t, x, got, want)
