"""Mesh preprocessing (ccfd_amd/csrc/mesh.c): what the mathematics and the
reference's construction fix, checked independently of the oracle."""
import os
import tempfile

import numpy as np
import pytest

from ccfd_amd import mesh as M
from tests import cases


def _per_slot_elem(m):
    return m.slot_elem()


def test_cartesian_closed_forms():
    """uniform quads: area dx*dy, axis-aligned unit normals, GP = half
    cell extent, LSQ weights of the central-difference stencil"""
    m = cases.sod_mesh(100, 2)
    dx, dy = 0.01, 0.005
    assert np.allclose(m.area, dx * dy, rtol=1e-12, atol=0)
    assert np.allclose(np.abs(m.n).max(axis=1), 1.0, rtol=0, atol=1e-15)
    assert np.allclose(np.abs(m.n).min(axis=1), 0.0, rtol=0, atol=1e-15)
    assert np.allclose(m.sx, dy, rtol=1e-12) and np.allclose(m.sy, dx, rtol=1e-12)
    assert abs(m.total_area_q * 0.01 - 1.0) < 1e-12
    # interior cell: w = d / (2 |d|^2) along each axis
    e = 150
    for s in range(m.slot_start[e], m.slot_start[e + 1]):
        d = m.bbv[s]
        ax = np.argmax(np.abs(d))
        assert abs(m.w[s, ax] - d[ax] / (2 * d[ax] ** 2)) < 1e-9 / abs(d[ax])
        assert abs(m.w[s, 1 - ax]) < 1e-6


@pytest.mark.parametrize("periodic", [True, False])
def test_delaunay_geometry_invariants(periodic):
    m = cases.delaunay_mesh(25, seed=5, periodic=periodic)
    se = _per_slot_elem(m)
    # closed cells: sum of n * len vanishes
    acc = np.zeros((m.n_elems, 2))
    np.add.at(acc, se, m.n * m.len[:, None])
    per = np.zeros(m.n_elems)
    np.add.at(per, se, m.len)
    assert (np.abs(acc).max(axis=1) / per).max() < 1e-13
    # outward normals
    assert ((m.n * m.GP).sum(axis=1) > 0).all()
    # total area is the unit square
    assert abs(m.area.sum() - 1.0) < 1e-12
    assert abs(1.0 / m.total_area_q - 1.0) < 1e-12
    # least squares: sum_s w_s (x) bbv_s = identity
    W = np.zeros((m.n_elems, 2, 2))
    np.add.at(W, se, m.w[:, :, None] * m.bbv[:, None, :])
    assert np.abs(W - np.eye(2)).max() < 1e-9
    # every non-ghost side has a partner facing back
    inner = m.nbr >= 0
    t = m.partner[inner]
    assert (m.partner[t] == np.nonzero(inner)[0]).all()
    assert np.array_equal(m.fn[inner], m.fn[t])
    if periodic:
        assert m.n_ghosts == 0
    else:
        assert m.n_ghosts == 4 * 24


def test_face_ownership_follows_creation_order():
    """after the stable (node0, node1, BC) sort the earlier-created side owns
    the face (src/mesh.c:1713-1755): the lower element index"""
    m = cases.delaunay_mesh(20, seed=1, periodic=False)
    se = _per_slot_elem(m)
    inner = np.nonzero(m.nbr >= 0)[0]
    own = inner[m.owner[inner] == 1]
    assert (se[own] < se[m.partner[own]]).all()
    # ghosts: the physical side owns
    assert (m.owner[m.nbr < 0] == 1).all()
    # one owner per face
    assert m.n_faces == (m.owner == 1).sum()


@pytest.mark.parametrize("hub", [0, 40])
def test_pairing_order_on_a_high_degree_node(hub):
    """a fan of 40 triangles around one node: with the hub numbered first its
    side-list bucket holds 80 entries (the sort's large-bucket branch), last
    it holds none. Either way every face pairs, the earlier-created side
    owns it and the boundary sides become ghosts in edge order."""
    k = 40
    ang = 2 * np.pi * np.arange(k) / k
    rim = np.stack([np.cos(ang), np.sin(ang)], axis=1)
    ids = [i for i in range(k + 1) if i != hub]   # rim node j -> ids[j]
    xy = np.zeros((k + 1, 2))
    xy[ids] = rim
    tri = np.array([[hub, ids[j], ids[(j + 1) % k], 0] for j in range(k)])
    bce = np.array([[ids[j], ids[(j + 1) % k], 201] for j in range(k)])
    m = M.Mesh(M.Raw(xy, tri=tri, bcedge=bce), [M.BC(201)])
    se = _per_slot_elem(m)
    inner = np.nonzero(m.nbr >= 0)[0]
    assert len(inner) == 2 * k
    assert (m.partner[m.partner[inner]] == inner).all()
    own = inner[m.owner[inner] == 1]
    assert len(own) == k and (se[own] < se[m.partner[own]]).all()
    # spoke j is shared by triangles j - 1 and j; triangle 0 owns spoke 0
    # (its neighbour is triangle k - 1, created later)
    assert sorted(se[own].tolist()) == sorted([0] + list(range(k - 1)))
    # the ghosts follow the sorted (node0, node1) order of the rim edges
    assert m.n_ghosts == k
    keys = [tuple(sorted((ids[j], ids[(j + 1) % k]))) for j in range(k)]
    want = [e for _, e in sorted(zip(keys, range(k)))]
    assert se[m.ghost_slot].tolist() == want
    assert abs(m.area.sum() - 0.5 * k * np.sin(2 * np.pi / k)) < 1e-13


def test_periodic_partners_translate():
    m = cases.delaunay_mesh(20, seed=1, periodic=True)
    se = _per_slot_elem(m)
    mid = m.bary[se] + m.GP
    midp = mid[m.partner]
    d = np.abs(mid - midp)
    wrap = (d > 0.5).any(axis=1)
    assert wrap.sum() == 4 * 19   # 19 edges on each of the 4 sides
    # translated by exactly one period in x or y
    dd = np.sort(d[wrap], axis=1)
    assert np.allclose(dd[:, 1], 1.0, atol=1e-12)
    assert np.allclose(dd[:, 0], 0.0, atol=1e-12)


def test_delaunay_empty_circumcircle():
    raw, _ = M.jittered_delaunay(15, 0.3, 9, periodic=False)
    xy, tri = raw.xy, raw.tri[:, :3]
    for t in tri[::7]:
        a, b, c = xy[t]
        d = 2 * (a[0] * (b[1] - c[1]) + b[0] * (c[1] - a[1]) + c[0] * (a[1] - b[1]))
        ux = ((a @ a) * (b[1] - c[1]) + (b @ b) * (c[1] - a[1]) + (c @ c) * (a[1] - b[1])) / d
        uy = ((a @ a) * (c[0] - b[0]) + (b @ b) * (a[0] - c[0]) + (c @ c) * (b[0] - a[0])) / d
        r2 = (a[0] - ux) ** 2 + (a[1] - uy) ** 2
        dist = ((xy - [ux, uy]) ** 2).sum(axis=1)
        dist[t] = np.inf
        assert (dist >= r2 * (1 - 1e-9)).all()


def test_gmsh_roundtrip():
    raw, bcs = M.jittered_delaunay(12, 0.3, 3, periodic=True)
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "m.msh")
        raw.write_gmsh(p)
        back = M.Raw.read_gmsh(p)
    assert np.array_equal(back.xy, raw.xy)
    assert np.array_equal(back.tri, raw.tri)
    assert np.array_equal(back.bcedge, raw.bcedge)
    m1, m2 = M.Mesh(raw, bcs), M.Mesh(back, bcs)
    assert np.array_equal(m1.w, m2.w) and np.array_equal(m1.nbr, m2.nbr)


def test_open_mesh_is_rejected():
    raw = M.cartesian(4, 4)
    raw.bcedge = raw.bcedge[:-1].copy()
    with pytest.raises(ValueError):
        M.Mesh(raw, [M.BC(101), M.BC(401)])


def test_undefined_bc_is_rejected():
    raw = M.cartesian(4, 4, side_bc=(101, 401, 101, 999))
    with pytest.raises(ValueError):
        M.Mesh(raw, [M.BC(101), M.BC(401)])


REF_CALC = "/root/reference/calc"
REF_MESHES = sorted(
    os.path.relpath(os.path.join(dp, f), REF_CALC)
    for dp, _, fs in os.walk(REF_CALC) for f in fs
    if f.endswith(".msh") or f.endswith(".mesh"))


def _bc_for(code):
    t = code // 100
    if t in (M.INFLOW, M.CHARACTERISTIC):
        return M.BC(code, rho=1.0, mach=0.5, alpha=0.0, pressure=1.0)
    if t == M.PRESSURE_OUT:
        return M.BC(code, pressure=1.0)
    return M.BC(code)


def _shoelace(xy, cells):
    x, y = xy[cells, 0], xy[cells, 1]
    return 0.5 * np.abs((x * np.roll(y, -1, axis=1) -
                         np.roll(x, -1, axis=1) * y).sum(axis=1))


@pytest.mark.parametrize("name", REF_MESHES or ["none"])
def test_reference_calc_meshes(name):
    """the meshes the reference ships with its cases (calc/: gmsh 2, gmsh 4
    and EMC2 files), read and preprocessed: closed cells with outward
    normals, the least-squares identity, partner sides facing back, every
    boundary edge one ghost side, and the total area of the cells computed
    independently"""
    if not REF_MESHES:
        pytest.skip("reference cases not present")
    path = os.path.join(REF_CALC, name)
    raw = M.Raw.read_emc2(path) if name.endswith(".mesh") else \
        M.Raw.read_gmsh(path)
    codes = sorted(set(raw.bcedge[:, 2].tolist()))
    m = M.Mesh(raw, [_bc_for(c) for c in codes])
    assert m.n_elems == len(raw.tri) + len(raw.quad)
    assert m.n_ghosts == len(raw.bcedge) and (m.nbr < 0).sum() == m.n_ghosts
    se = _per_slot_elem(m)
    acc = np.zeros((m.n_elems, 2))
    np.add.at(acc, se, m.n * m.len[:, None])
    per = np.zeros(m.n_elems)
    np.add.at(per, se, m.len)
    assert (np.abs(acc).max(axis=1) / per).max() < 1e-12
    assert ((m.n * m.GP).sum(axis=1) > 0).all()
    assert (m.area > 0).all()
    total = _shoelace(raw.xy, raw.tri[:, :3]).sum() + \
        _shoelace(raw.xy, raw.quad[:, :4]).sum()
    assert abs(m.area.sum() / total - 1.0) < 1e-12
    W = np.zeros((m.n_elems, 2, 2))
    np.add.at(W, se, m.w[:, :, None] * m.bbv[:, None, :])
    assert np.abs(W - np.eye(2)).max() < 1e-8
    inner = m.nbr >= 0
    t = m.partner[inner]
    assert (m.partner[t] == np.nonzero(inner)[0]).all()
    assert np.allclose(m.n[inner], -m.n[t], rtol=0, atol=1e-14)


@pytest.mark.parametrize("bad,match", [
    ("node", "references node"), ("negative", "references node"),
    ("degenerate", "unmatched side"), ("bc_node", "missing node"),
    ("bc_twice", "open mesh")])
def test_malformed_input_is_rejected(bad, match):
    """bad connectivity fails with a message instead of reading out of
    bounds"""
    raw = M.cartesian(4, 4)
    raw.quad, raw.bcedge = raw.quad.copy(), raw.bcedge.copy()
    if bad == "node":
        raw.quad[0, 0] = 10 ** 6
    elif bad == "negative":
        raw.quad[0, 1] = -5
    elif bad == "degenerate":
        raw.quad[0, 1] = raw.quad[0, 0]
    elif bad == "bc_node":
        raw.bcedge[0, 0] = 10 ** 6
    else:
        raw.bcedge = np.vstack([raw.bcedge, raw.bcedge[:1]])
    with pytest.raises(ValueError, match=match):
        M.Mesh(raw, [M.BC(101), M.BC(401)])


def test_element_orientation_does_not_matter():
    """cells listed clockwise: the reference takes |cross products| for the
    area and turns every normal outward (src/mesh.c:250-341), so the same
    areas and outward normals come out; and where the reference is built,
    its fvTimeDerivative and the oracle's agree bit for bit on such a mesh"""
    raw, bcs = M.jittered_delaunay(10, 0.3, 4, periodic=False)
    m0 = M.Mesh(raw, bcs)
    raw.tri = raw.tri.copy()
    raw.tri[::2, :3] = raw.tri[::2, 2::-1]
    m1 = M.Mesh(raw, bcs)
    assert np.allclose(m1.area, m0.area, rtol=1e-14, atol=0)
    assert ((m1.n * m1.GP).sum(axis=1) > 0).all()
    se = _per_slot_elem(m1)
    acc = np.zeros((m1.n_elems, 2))
    np.add.at(acc, se, m1.n * m1.len[:, None])
    assert np.abs(acc).max() < 1e-13
    from oracle import bindings as B
    if not B.ref_available():
        return
    prim = cases.random_state(m1.n_elems, seed=3, mach=0.5)
    ini = os.path.join(tempfile.mkdtemp(), "flip.ini")
    B.write_ini(ini, [("fluxFunction", 2), ("nBC", 1), ("BCtype", 401),
                      ("spatialOrder", 2), ("limiter", 1), ("venk_k", 1.0)])
    ref = B.Ref()
    ref.setup(ini, m1)
    ref.set_prim(prim)
    r = ref.fvtd(0.0)
    import ctypes
    ctypes.CDLL(None).fflush(None)
    o = B.Oracle(m1, B.make_cfg(flux=2, order=2, limiter=1))
    o.set_prim(prim)
    assert np.array_equal(o.fvtd(0.0), r["u_t"])
