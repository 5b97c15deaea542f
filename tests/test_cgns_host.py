"""CGNS output and restart files (include/ccfd_cgns.h, libccfd_cgns.so) on
the host.

The reference writes these through the CGNS library (src/output.c:122-244,
407-517, 532-668) and reads a restart with cgnsReadSolution
(src/initialCondition.c:94-158). That library is not in this image, so the
files are checked against the CGNS/HDF5 node layout itself: every node's
name, label, data type, flags, dimensions (stored in reverse order) and
values, in the order the reference's calls create them, dumped by a small
HDF5 walker (tests/native/cgns_tree.c); links are external links with their
file and path. The expected trees are built here from the mesh arrays
directly (a link's file and path are stored with their terminating NUL,
shown as "~"). The reader is checked by round trips and refusals."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from ccfd_amd import _native as N
from ccfd_amd import build as BLD
from ccfd_amd import mesh as M
from tests import cases

LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "ccfd_amd", "lib", "libccfd_cgns.so")
TREE_SRC = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native",
                        "cgns_tree.c")
EDIT_SRC = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native",
                        "cgns_edit.c")


@pytest.fixture(scope="module")
def cg():
    pre = BLD.hdf5_prefix()
    if pre is None or not os.path.exists(LIB):
        pytest.skip("HDF5 (and libccfd_cgns.so) not available")
    L = C.CDLL(LIB)
    vp, cp, dp, lp = C.c_void_p, C.c_char_p, N.c_double_p, N.c_long_p
    L.ccfd_cgns_write_mesh.argtypes = [cp, vp, C.c_long, lp, cp]
    L.ccfd_cgns_write_solution.argtypes = [cp, cp, vp, dp, C.c_double,
                                           C.c_double, cp]
    L.ccfd_cgns_write_master.argtypes = [cp, cp, vp, C.c_int, vp, vp, vp, cp]
    L.ccfd_cgns_read_solution.argtypes = [cp, C.c_long, dp,
                                          C.POINTER(C.c_double),
                                          C.POINTER(C.c_double), cp]
    PL, PD = C.POINTER(C.c_long), C.POINTER(C.POINTER(C.c_double))
    PPL = C.POINTER(C.POINTER(C.c_long))
    L.ccfd_cgns_read_mesh.argtypes = [cp, PL, PD, PL, PPL, PL, PPL, PL, PPL,
                                      cp]
    return L, pre


@pytest.fixture(scope="module")
def tree(cg, tmp_path_factory):
    _, pre = cg
    exe = str(tmp_path_factory.mktemp("tree") / "cgns_tree")
    r = subprocess.run(["gcc", "-std=c99", "-O1", "-I" + os.path.join(pre, "include"),
                        TREE_SRC, "-o", exe, "-L" + os.path.join(pre, "lib"),
                        "-lhdf5", "-Wl,-rpath," + os.path.join(pre, "lib")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr

    def dump(path):
        out = subprocess.run([exe, path], capture_output=True, text=True,
                             check=True).stdout
        nodes, order = {}, []
        for line in out.splitlines():
            f = line.split("|")
            if f[1].startswith("@"):
                nodes[f[1]] = f[2]
                continue
            nodes[f[0]] = {"name": f[1], "label": f[2], "type": f[3],
                           "flags": f[4], "dims": f[5], "data": f[6]}
            order.append(f[0])
        return nodes, order
    return dump


@pytest.fixture(scope="module")
def edit(cg, tmp_path_factory):
    _, pre = cg
    exe = str(tmp_path_factory.mktemp("edit") / "cgns_edit")
    r = subprocess.run(["gcc", "-std=c99", "-O1", "-I" + os.path.join(pre, "include"),
                        EDIT_SRC, "-o", exe, "-L" + os.path.join(pre, "lib"),
                        "-lhdf5", "-Wl,-rpath," + os.path.join(pre, "lib")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr

    def run(*args):
        subprocess.run([exe] + [str(a) for a in args], check=True)
    return run


def _read_mesh(L, path):
    """ccfd_cgns_read_mesh: (Raw, None) or (None, message)"""
    n = [C.c_long() for _ in range(4)]
    xy = C.POINTER(C.c_double)()
    arrs = [C.POINTER(C.c_long)() for _ in range(3)]
    err = C.create_string_buffer(256)
    rc = L.ccfd_cgns_read_mesh(str(path).encode(), C.byref(n[0]), C.byref(xy),
                               C.byref(n[1]), C.byref(arrs[0]), C.byref(n[2]),
                               C.byref(arrs[1]), C.byref(n[3]),
                               C.byref(arrs[2]), err)
    if rc:
        return None, err.value.decode()
    libc = C.CDLL(None)
    libc.free.argtypes = [C.c_void_p]

    def take(p, count):
        a = (np.ctypeslib.as_array(p, shape=(count,)).copy() if count
             else np.zeros(0))
        libc.free(C.cast(p, C.c_void_p))
        return a
    return M.Raw(take(xy, 2 * n[0].value), take(arrs[0], 4 * n[1].value),
                 take(arrs[1], 5 * n[2].value),
                 take(arrs[2], 3 * n[3].value)), None


def _mesh():
    raw = cases.split_quads(M.cartesian(5, 3, side_bc=(101, 401, 301, 401)),
                            seed=2)
    bcs = [M.BC(401), M.BC(101),
           M.BC(301, rho=1.0, mach=0.3, alpha=0.0, pressure=1.0)]
    return raw, bcs, M.Mesh(raw, bcs)


def _ints(text):
    return [int(float(x)) for x in text.split(",")] if text else []


def _node(nodes, path, label, typ, dims, data=None):
    n = nodes[path]
    assert (n["label"], n["type"], n["flags"]) == (label, typ, "1"), (path, n)
    assert n["name"] == path.rsplit("/", 1)[-1]
    assert n["dims"] == dims, (path, n["dims"])
    if data is not None:
        assert n["data"] == data, (path, n["data"][:80])


def _expected_bars(raw, m):
    """Boundaries (src/output.c:576-601): per boundary condition in
    parameter order, the boundary sides in the reference's list order (the
    list is built by prepending, so the last ghost first), each with the
    mesh file's edge nodes, numbered from 1; and the element partition"""
    key = {}
    for a, b, _ in raw.bcedge.tolist():
        key.setdefault((min(a, b), max(a, b)), (a, b))
    se = m.slot_node
    bars, part = [], [0]
    for c in range(len(m.bcs)):
        for g in range(m.n_ghosts - 1, -1, -1):
            if m.ghost_bc[g] != c:
                continue
            s = m.ghost_slot[g]
            a, b = int(se[s, 0]), int(se[s, 1])
            bars += [x + 1 for x in key[(min(a, b), max(a, b))]]
        part.append(len(bars) // 2)
    return bars, part


def test_grid_file_layout(cg, tree, tmp_path):
    L, _ = cg
    raw, bcs, m = _mesh()
    path = str(tmp_path / "case_mesh.cgns")
    err = C.create_string_buffer(256)
    assert L.ccfd_cgns_write_mesh(path.encode(), C.cast(m._p, C.c_void_p),
                                  len(raw.bcedge), N.lptr(raw.bcedge),
                                  err) == 0, err.value
    nodes, order = tree(path)
    root = nodes["/"]
    assert (root["name"], root["label"], root["type"]) == (
        "HDF5 MotherNode", "Root Node of HDF5 File", "MT")
    assert nodes["@format"] == "IEEE_LITTLE_32~"
    assert nodes["@hdf5version"].startswith("HDF5 Version 1.")
    assert len(nodes["@hdf5version"]) == 33
    nV, nE = len(raw.xy), m.n_elems
    nT, nQ, nG = len(raw.tri), len(raw.quad), m.n_ghosts
    _node(nodes, "/CGNSLibraryVersion", "CGNSLibraryVersion_t", "R4", "1",
          "%.17g" % np.float32(4.1))
    _node(nodes, "/Base", "CGNSBase_t", "I4", "2", "2,3")
    # unstructured zone: [IndexDimension = 1][3], stored reversed
    _node(nodes, "/Base/Zone", "Zone_t", "I4", "3x1", "%d,%d,0" % (nV, nE))
    _node(nodes, "/Base/Zone/ZoneType", "ZoneType_t", "C1", "12", "Unstructured")
    _node(nodes, "/Base/Zone/GridCoordinates", "GridCoordinates_t", "MT", "")
    for k, c in enumerate("XYZ"):
        want = raw.xy[:, k] if k < 2 else np.zeros(nV)
        _node(nodes, "/Base/Zone/GridCoordinates/Coordinate" + c,
              "DataArray_t", "R8", str(nV),
              ",".join("%.17g" % v for v in want))
    for name, typ, lo, conn in (("Triangles", 5, 1, raw.tri[:, :3]),
                                ("Quadrilaterals", 7, nT + 1, raw.quad[:, :4])):
        p = "/Base/Zone/" + name
        _node(nodes, p, "Elements_t", "I4", "2", "%d,0" % typ)
        _node(nodes, p + "/ElementRange", "IndexRange_t", "I4", "2",
              "%d,%d" % (lo, lo + len(conn) - 1))
        _node(nodes, p + "/ElementConnectivity", "DataArray_t", "I4",
              str(conn.size), ",".join(str(x + 1) for x in conn.ravel()))
    bars, part = _expected_bars(raw, m)
    assert len(bars) == 2 * nG
    p = "/Base/Zone/Boundaries"
    _node(nodes, p, "Elements_t", "I4", "2", "3,0")
    _node(nodes, p + "/ElementRange", "IndexRange_t", "I4", "2",
          "%d,%d" % (nE + 1, nE + nG))
    assert _ints(nodes[p + "/ElementConnectivity"]["data"]) == bars
    _node(nodes, "/Base/Zone/ZoneBC", "ZoneBC_t", "MT", "")
    for c, b in enumerate(bcs):
        q = "/Base/Zone/ZoneBC/%d" % b.code
        _node(nodes, q, "BC_t", "C1", "9", "BCGeneral")
        n = part[c + 1] - part[c]
        # ElementList is stored as a PointList of edge elements (cell
        # dimension 2): [IndexDimension][n], reversed
        _node(nodes, q + "/PointList", "IndexArray_t", "I4", "%dx1" % n,
              ",".join(str(nE + part[c] + i + 1) for i in range(n)))
        _node(nodes, q + "/GridLocation", "GridLocation_t", "C1", "10",
              "EdgeCenter")
    # creation order: the reference's call order
    top = [p for p in order if p.count("/") <= 3]
    assert top == ["/", "/CGNSLibraryVersion", "/Base", "/Base/Zone",
                   "/Base/Zone/ZoneType", "/Base/Zone/GridCoordinates",
                   "/Base/Zone/Triangles", "/Base/Zone/Quadrilaterals",
                   "/Base/Zone/Boundaries", "/Base/Zone/ZoneBC"]


def test_periodic_grid_has_no_boundary_section(cg, tree, tmp_path):
    L, _ = cg
    raw, bcs = M.jittered_delaunay(6, 0.3, 3, periodic=True)
    m = M.Mesh(raw, bcs)
    path = str(tmp_path / "per_mesh.cgns")
    err = C.create_string_buffer(256)
    assert L.ccfd_cgns_write_mesh(path.encode(), C.cast(m._p, C.c_void_p),
                                  len(raw.bcedge), N.lptr(raw.bcedge), err) == 0
    nodes, _ = tree(path)
    assert "/Base/Zone/Triangles" in nodes
    assert "/Base/Zone/Boundaries" not in nodes and "/Base/Zone/ZoneBC" not in nodes


def _write_solution(L, m, path, prim, t, tover=0.0, grid="case_mesh.cgns"):
    err = C.create_string_buffer(256)
    rc = L.ccfd_cgns_write_solution(path.encode(), grid.encode(),
                                    C.cast(m._p, C.c_void_p), N.dptr(prim),
                                    t, tover, err)
    assert rc == 0, err.value


def test_solution_file_and_restart_read(cg, tree, tmp_path):
    L, _ = cg
    raw, bcs, m = _mesh()
    rng = np.random.default_rng(7)
    prim = np.ascontiguousarray(rng.uniform(0.2, 3.0, (m.n_elems, 4)))
    path = str(tmp_path / "case_0000000.1250000.cgns")
    _write_solution(L, m, path, prim, 0.125)
    nodes, order = tree(path)
    for name in ("GridCoordinates", "Triangles", "Quadrilaterals"):
        _node(nodes, "/Base/Zone/" + name, "", "LK", "",
              "case_mesh.cgns~:/Base/Zone/%s~:external" % name)
    fs = "/Base/Zone/FlowSolution"
    _node(nodes, fs, "FlowSolution_t", "MT", "")
    _node(nodes, fs + "/GridLocation", "GridLocation_t", "C1", "10", "CellCenter")
    cols = {"Density": prim[:, 0], "VelocityX": prim[:, 1],
            "VelocityY": prim[:, 2], "VelocityZ": np.zeros(m.n_elems),
            "Pressure": prim[:, 3]}
    for name, v in cols.items():
        _node(nodes, fs + "/" + name, "DataArray_t", "R8", str(m.n_elems),
              ",".join("%.17g" % x for x in v))
    _node(nodes, "/Base/ConvergenceInfo", "Descriptor_t", "C1", "41",
          "%20.12f %20.12f" % (0.125, 0.0))
    assert order.index("/Base/ConvergenceInfo") > order.index(fs + "/Pressure")
    # restart read: the state bit for bit, the time from the descriptor
    out = np.zeros_like(prim)
    t, to = C.c_double(), C.c_double()
    err = C.create_string_buffer(256)
    assert L.ccfd_cgns_read_solution(path.encode(), m.n_elems, N.dptr(out),
                                     C.byref(t), C.byref(to), err) == 0, err.value
    assert np.array_equal(out, prim) and t.value == 0.125 and to.value == 0.0
    # the descriptor keeps 12 decimals, as the reference's "%20.12f"
    p2 = str(tmp_path / "case_2.cgns")
    _write_solution(L, m, p2, prim, 1.0 / 3.0, 2.5)
    assert L.ccfd_cgns_read_solution(p2.encode(), m.n_elems, N.dptr(out),
                                     C.byref(t), C.byref(to), err) == 0
    assert t.value == float("%.12f" % (1.0 / 3.0)) and to.value == 2.5
    # refusals
    assert L.ccfd_cgns_read_solution(path.encode(), m.n_elems + 1, N.dptr(out),
                                     C.byref(t), C.byref(to), err) == -1
    assert b"Wrong Number of Elements" in err.value
    assert L.ccfd_cgns_read_solution(str(tmp_path / "none.cgns").encode(),
                                     m.n_elems, N.dptr(out), C.byref(t),
                                     C.byref(to), err) == -1
    assert b"cannot open" in err.value


def test_master_file(cg, tree, tmp_path):
    L, _ = cg
    raw, bcs, m = _mesh()
    iters = (C.c_long * 3)(20, 40, 60)
    times = (C.c_double * 3)(0.5, 1.0, 1.5)
    names = [b"case_%09d.cgns" % i for i in (20, 40, 60)]
    files = (C.c_char_p * 3)(*names)
    path = str(tmp_path / "case_Master.cgns")
    err = C.create_string_buffer(256)
    assert L.ccfd_cgns_write_master(path.encode(), b"case_mesh.cgns",
                                    C.cast(m._p, C.c_void_p), 3, times, iters,
                                    files, err) == 0, err.value
    nodes, order = tree(path)
    for k, i in enumerate((20, 40, 60)):
        _node(nodes, "/Base/Zone/FlowSolution%09d" % i, "", "LK", "",
              "case_%09d.cgns~:/Base/Zone/FlowSolution~:external" % i)
    _node(nodes, "/Base/TimeIterValues", "BaseIterativeData_t", "I4", "1", "3")
    _node(nodes, "/Base/TimeIterValues/TimeValues", "DataArray_t", "R8", "3",
          "0.5,1,1.5")
    # the reference passes its long iteration numbers as Integer: the
    # array holds their first three 32-bit words
    _node(nodes, "/Base/TimeIterValues/IterationValues", "DataArray_t", "I4",
          "3", "20,0,40")
    _node(nodes, "/Base/Zone/ZoneIterativeData", "ZoneIterativeData_t", "MT", "")
    ptr = "".join("FlowSolution%09d" % i + "~" * 11 for i in (20, 40, 60))
    _node(nodes, "/Base/Zone/ZoneIterativeData/FlowSolutionPointers",
          "DataArray_t", "C1", "3x32", ptr)
    assert order.index("/Base/TimeIterValues") > order.index(
        "/Base/Zone/FlowSolution000000060")
    # an output listed twice (a periodic output at the last iteration, then
    # the final one) is linked once and listed twice
    iters = (C.c_long * 3)(20, 40, 40)
    files = (C.c_char_p * 3)(names[0], names[1], names[1])
    assert L.ccfd_cgns_write_master(path.encode(), b"case_mesh.cgns",
                                    C.cast(m._p, C.c_void_p), 3, times, iters,
                                    files, err) == 0, err.value
    nodes, order = tree(path)
    assert [p for p in order if "/FlowSolution0" in p] == [
        "/Base/Zone/FlowSolution000000020", "/Base/Zone/FlowSolution000000040"]
    assert nodes["/Base/Zone/ZoneIterativeData/FlowSolutionPointers"][
        "dims"] == "3x32"


def test_null_arguments_are_refused(cg):
    L, _ = cg
    err = C.create_string_buffer(256)
    assert L.ccfd_cgns_write_mesh(None, None, 0, None, err) == -1
    assert err.value == b"null argument"
    assert L.ccfd_cgns_write_solution(b"x.cgns", None, None, None, 0.0, 0.0,
                                      err) == -1
    assert L.ccfd_cgns_write_master(b"x.cgns", b"g.cgns", None, 0, None, None,
                                    None, err) == -1
    t = C.c_double()
    assert L.ccfd_cgns_read_solution(b"x.cgns", 0, None, C.byref(t), C.byref(t),
                                     err) == -1


def _written(L, tmp_path, name="m.cgns"):
    raw, bcs, m = _mesh()
    path = tmp_path / name
    err = C.create_string_buffer(256)
    assert L.ccfd_cgns_write_mesh(str(path).encode(), C.cast(m._p, C.c_void_p),
                                  len(raw.bcedge), N.lptr(raw.bcedge),
                                  err) == 0, err.value
    return raw, bcs, m, path


def test_mesh_file_round_trip(cg, tmp_path):
    """readCGNS (src/mesh.c:1153-1394) of the grid file cgnsWriteMesh
    writes: the nodes, the cells in element order (domain 1), and the
    boundary edges of the Boundaries section with the codes of the BC
    nodes in order; the mesh built from it equals the original"""
    L, _ = cg
    raw, bcs, m, path = _written(L, tmp_path)
    got, msg = _read_mesh(L, path)
    assert msg is None, msg
    assert np.array_equal(got.xy, raw.xy)
    assert np.array_equal(got.tri, raw.tri)
    assert np.array_equal(got.quad, raw.quad)
    bars, part = _expected_bars(raw, m)
    want = []
    for c, bc in enumerate(bcs):
        for i in range(part[c], part[c + 1]):
            want.append([bars[2 * i] - 1, bars[2 * i + 1] - 1, bc.code])
    assert got.bcedge.tolist() == want
    m2 = M.Mesh(got, bcs)
    assert m2.n_elems == m.n_elems and m2.n_ghosts == m.n_ghosts
    for a in ("elem_node", "bary", "area", "sx", "sy", "slot_start"):
        assert np.array_equal(getattr(m2, a), getattr(m, a)), a


def test_mesh_reader_follows_the_library_and_the_reference(cg, edit, tmp_path):
    """the reference's reading, with the CGNS 4.1 library it vendors: BC
    codes go to consecutive boundary edges, as many per BC as its point set
    counts (a range counts 2), whatever the set holds; zone 1 is the first
    zone by name; MIXED sections are refused by the library, other element
    types by the reference; a BC name must be a code"""
    L, _ = cg
    raw, bcs, m, path = _written(L, tmp_path)
    n_side = {bc.code: int((raw.bcedge[:, 2] == bc.code).sum()) for bc in bcs}
    # PointRange on the first BC (401, six edges)
    p = tmp_path / "range.cgns"
    p.write_bytes(path.read_bytes())
    edit("range", p, "/Base/Zone/ZoneBC/401", 41, 46)
    got, msg = _read_mesh(L, p)
    assert msg is None, msg
    codes = [401] * 2 + [101] * n_side[101] + [301] * n_side[301]
    codes += [0] * (len(got.bcedge) - len(codes))
    assert got.bcedge[:, 2].tolist() == codes
    # a second zone that sorts first, without the quadrilaterals
    p = tmp_path / "zones.cgns"
    p.write_bytes(path.read_bytes())
    edit("copy", p, "/Base/Zone", "/Base/Abc")
    edit("delete", p, "/Base/Abc/Quadrilaterals")
    got, msg = _read_mesh(L, p)
    assert msg is None, msg
    assert len(got.quad) == 0 and np.array_equal(got.tri, raw.tri)
    # refusals
    for name, args, text in [
            ("mixed", ("settype", "/Base/Zone/Triangles", 20), "MIXED"),
            ("tri6", ("settype", "/Base/Zone/Triangles", 6),
             "Only TRI_3 and QUAD_4 Elements allowed"),
            ("name", ("rename", "/Base/Zone/ZoneBC/101", "wall"),
             "boundary condition wall is not a boundary code"),
            ("extra", ("copy", "/Base/Zone/ZoneBC/401", "/Base/Zone/ZoneBC/402"),
             "more boundary condition points than boundary edges")]:
        p = tmp_path / (name + ".cgns")
        p.write_bytes(path.read_bytes())
        edit(args[0], p, *args[1:])
        got, msg = _read_mesh(L, p)
        assert got is None and text in msg, (name, msg)
    got, msg = _read_mesh(L, tmp_path / "missing.cgns")
    assert got is None and "cannot open" in msg
