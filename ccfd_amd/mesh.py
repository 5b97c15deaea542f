"""Meshes: raw connectivity generators and the preprocessed host mesh.

`Mesh` wraps the C preprocessing of include/ccfd_mesh.h (reference element
order, face ownership and geometry bit-identical to ccfd's src/mesh.c).
Generators produce raw connectivity (vertices, triangles, quadrilaterals,
boundary edges) for the workloads of this path:

  cartesian        - createCartMesh (src/mesh.c:385-486), same vertex
                     accumulation as the reference
  jittered_delaunay- Delaunay triangulation of a seeded jittered grid on the
                     unit square, optionally periodic in x and y
  naca_ogrid       - body-fitted O-grid around an analytic NACA 4-digit
                     section (closed trailing edge)
  channel_bump     - channel with a circular-arc bump on the lower wall
"""
import ctypes as C
import math

import numpy as np

from . import _native as N

# boundary types, ccfd src/main.h:44-53
SLIPWALL, WALL, INFLOW, OUTFLOW, CHARACTERISTIC, EXACTSOL, PERIODIC, \
    PRESSURE_OUT = range(1, 9)


class BC:
    """One boundary condition (initBoundary, src/boundary.c:29-137).

    code = 100*type + id. For INFLOW / CHARACTERISTIC give rho, mach, alpha
    (deg), pressure; PRESSURE_OUT needs pressure; PERIODIC needs connection;
    EXACTSOL needs exact_func.
    """

    def __init__(self, code, rho=None, mach=None, alpha=None, pressure=None,
                 connection=(0.0, 0.0), exact_func=0, gamma=1.4, pvar=None):
        self.code = int(code)
        self.type = self.code // 100
        self.id = self.code % 100
        self.exact_func = exact_func
        self.connection = tuple(connection)
        self.pvar = [0.0, 0.0, 0.0, 0.0]
        if pvar is not None:
            self.pvar = list(pvar)
        elif self.type in (INFLOW, CHARACTERISTIC):
            # same libm as the reference (src/boundary.c:83-96)
            pi = math.acos(-1.0)
            c = math.sqrt(gamma * pressure / rho)
            v = mach * c
            self.pvar = [rho, v * math.cos(alpha * pi / 180.0),
                         v * math.sin(alpha * pi / 180.0), pressure]
        elif self.type == PRESSURE_OUT:
            self.pvar = [0.0, 0.0, 0.0, pressure]

    def to_c(self):
        b = N.CcfdBC()
        b.type, b.id, b.exactFunc = self.type, self.id, int(self.exact_func)
        for i in range(4):
            b.pVar[i] = float(self.pvar[i])
        b.connection[0], b.connection[1] = self.connection
        return b


class Raw:
    """Raw connectivity: xy (n,2), tri (nt,4: nodes+domain), quad (nq,5),
    bcedge (nb,3: node0, node1, code)."""

    def __init__(self, xy, tri=None, quad=None, bcedge=None):
        self.xy = np.ascontiguousarray(xy, dtype=np.float64).reshape(-1, 2)
        self.tri = np.ascontiguousarray(
            tri if tri is not None else np.zeros((0, 4)), dtype=np.int64
        ).reshape(-1, 4)
        self.quad = np.ascontiguousarray(
            quad if quad is not None else np.zeros((0, 5)), dtype=np.int64
        ).reshape(-1, 5)
        self.bcedge = np.ascontiguousarray(
            bcedge if bcedge is not None else np.zeros((0, 3)), dtype=np.int64
        ).reshape(-1, 3)

    def write_gmsh(self, path):
        rc = N.host().ccfd_gmsh_write(
            path.encode(), len(self.xy), N.dptr(self.xy), len(self.tri),
            N.lptr(self.tri), len(self.quad), N.lptr(self.quad),
            len(self.bcedge), N.lptr(self.bcedge))
        if rc != 0:
            raise IOError("could not write " + path)

    @staticmethod
    def read_gmsh(path):
        """gmsh ASCII, format 2 or 4"""
        return Raw._read(N.host().ccfd_gmsh_read, path)

    @staticmethod
    def read_emc2(path):
        """EMC2 / medit .mesh"""
        return Raw._read(N.host().ccfd_emc2_read, path)

    @staticmethod
    def _read(fn, path):
        nv, nt, nq, nb = C.c_long(), C.c_long(), C.c_long(), C.c_long()
        xy, tri, quad, bce = N.c_double_p(), N.c_long_p(), N.c_long_p(), \
            N.c_long_p()
        err = C.create_string_buffer(256)
        rc = fn(path.encode(), C.byref(nv), C.byref(xy), C.byref(nt),
                C.byref(tri), C.byref(nq), C.byref(quad), C.byref(nb),
                C.byref(bce), err)
        if rc != 0:
            raise IOError(err.value.decode())
        r = Raw(N.view(xy, 2 * nv.value, np.float64),
                N.view(tri, 4 * nt.value, np.int64),
                N.view(quad, 5 * nq.value, np.int64),
                N.view(bce, 3 * nb.value, np.int64))
        for p in (xy, tri, quad, bce):
            N.libc.free(C.cast(p, C.c_void_p))
        return r


class RawMesh:
    """Raw connectivity and boundary conditions handed to the device as they
    are (struct ccfd_raw_mesh): the batched arena preprocesses it on the GPU
    at admission (ccfd_batch_add_raw). Keeps the arrays alive while the
    struct points at them."""

    def __init__(self, raw, bcs):
        self.raw = raw
        self.bcs = list(bcs)
        self.n_elems = len(raw.tri) + len(raw.quad)
        self._bc = (N.CcfdBC * max(1, len(self.bcs)))()
        for i, b in enumerate(self.bcs):
            self._bc[i] = b.to_c()
        self._s = N.CcfdRawMesh(
            len(raw.xy), N.dptr(raw.xy), len(raw.tri), N.lptr(raw.tri),
            len(raw.quad), N.lptr(raw.quad), len(raw.bcedge),
            N.lptr(raw.bcedge), len(self.bcs), self._bc)
        self.ptr = C.pointer(self._s)


class Mesh:
    """Preprocessed mesh in the reference's element / slot order."""

    def __init__(self, raw, bcs):
        self.raw = raw
        self.bcs = list(bcs)
        lib = N.host()
        arr = (N.CcfdBC * max(1, len(self.bcs)))()
        for i, b in enumerate(self.bcs):
            arr[i] = b.to_c()
        err = C.create_string_buffer(256)
        p = lib.ccfd_mesh_build(
            len(raw.xy), N.dptr(raw.xy), len(raw.tri), N.lptr(raw.tri),
            len(raw.quad), N.lptr(raw.quad), len(raw.bcedge),
            N.lptr(raw.bcedge), len(self.bcs), arr, err)
        if not p:
            raise ValueError("mesh build failed: " + err.value.decode())
        self._p = p
        m = p.contents
        self.n_elems = m.nElems
        self.n_slots = m.nSlots
        self.n_faces = m.nFaces
        self.n_ghosts = m.nGhosts
        self.total_area_q = m.totalArea_q
        self.bbox = (m.xMin, m.xMax, m.yMin, m.yMax)
        self._cache = {}

    # name -> (struct field, count kind, per-item width, dtype); arrays are
    # copied out of the C structure on first access
    _ARRAYS = {
        "elem_type": ("elemType", "e", 1, np.int32),
        "elem_node": ("elemNode", "e", 4, np.int64),
        "bary": ("bary", "e", 2, np.float64),
        "area": ("area", "e", 1, np.float64),
        "areaq": ("areaq", "e", 1, np.float64),
        "sx": ("sx", "e", 1, np.float64),
        "sy": ("sy", "e", 1, np.float64),
        "slot_start": ("slotStart", "e+1", 1, np.int64),
        "nbr": ("nbr", "s", 1, np.int64),
        "partner": ("partner", "s", 1, np.int64),
        "owner": ("owner", "s", 1, np.uint8),
        "n": ("n", "s", 2, np.float64),
        "len": ("len", "s", 1, np.float64),
        "fn": ("fn", "s", 2, np.float64),
        "flen": ("flen", "s", 1, np.float64),
        "GP": ("GP", "s", 2, np.float64),
        "GPc": ("GPc", "s", 2, np.float64),
        "w": ("w", "s", 2, np.float64),
        "bbv": ("bbv", "s", 2, np.float64),
        "bbd": ("bbd", "s", 1, np.float64),
        "slot_node": ("slotNode", "s", 2, np.int64),
        "ghost_slot": ("ghostSlot", "g", 1, np.int64),
        "ghost_bc": ("ghostBC", "g", 1, np.int32),
        "ghost_bary": ("ghostBary", "g", 2, np.float64),
        "face_slot": ("faceSlot", "f", 1, np.int64),
    }

    def __getattr__(self, name):
        spec = Mesh._ARRAYS.get(name)
        if spec is None or "_p" not in self.__dict__:
            raise AttributeError(name)
        cache = self.__dict__["_cache"]
        if name not in cache:
            field, kind, width, dt = spec
            n = {"e": self.n_elems, "e+1": self.n_elems + 1,
                 "s": self.n_slots, "g": self.n_ghosts,
                 "f": self.n_faces}[kind]
            a = N.view(getattr(self._p.contents, field), n * width, dt)
            cache[name] = a.reshape(n, width) if width > 1 else a
        return cache[name]

    @property
    def ptr(self):
        return self._p

    def slot_elem(self):
        return np.repeat(np.arange(self.n_elems),
                         np.diff(self.slot_start)).astype(np.int64)

    def __del__(self):
        p = getattr(self, "_p", None)
        if p:
            try:
                N.host().ccfd_mesh_free(p)
            except Exception:
                pass
            self._p = None


# ---------------------------------------------------------------- generators

def cartesian(i_max, j_max, x0=(0.0, 0.0), x1=(1.0, 1.0),
              side_bc=(101, 401, 101, 401)):
    """createCartMesh (src/mesh.c:385-486) with one BC segment per side
    (bottom, right, top, left)."""
    lib = N.host()
    nseg = np.ones(4, dtype=np.int32)
    btype = np.zeros((4, 20), dtype=np.int32)
    brange = np.zeros((4, 20, 2), dtype=np.int32)
    for s in range(4):
        btype[s, 0] = side_bc[s]
        brange[s, 0] = (1, i_max if s % 2 == 0 else j_max)
    nv, nq, nb = C.c_long(), C.c_long(), C.c_long()
    xy, quad, bce = N.c_double_p(), N.c_long_p(), N.c_long_p()
    rc = lib.ccfd_cart_connectivity(
        i_max, j_max, x0[0], x0[1], x1[0], x1[1], N.iptr(nseg),
        N.iptr(btype.reshape(-1)), N.iptr(brange.reshape(-1)), C.byref(nv),
        C.byref(xy), C.byref(nq), C.byref(quad), C.byref(nb), C.byref(bce))
    if rc != 0:
        raise MemoryError("cartesian mesh")
    raw = Raw(N.view(xy, 2 * nv.value, np.float64), None,
              N.view(quad, 5 * nq.value, np.int64),
              N.view(bce, 3 * nb.value, np.int64))
    for p in (xy, quad, bce):
        N.libc.free(C.cast(p, C.c_void_p))
    return raw


def jittered_points(n, jitter=0.3, seed=2):
    """n x n grid on the unit square, interior points jittered by
    U(-jitter h, jitter h) per coordinate (PCG64 with the given seed);
    boundary points stay on the grid."""
    h = 1.0 / (n - 1)
    g = np.arange(n, dtype=np.float64) * h
    g[-1] = 1.0
    X, Y = np.meshgrid(g, g, indexing="xy")
    rng = np.random.Generator(np.random.PCG64(seed))
    d = rng.uniform(-jitter * h, jitter * h, size=(2, n, n))
    inner = np.zeros((n, n), dtype=bool)
    inner[1:-1, 1:-1] = True
    X = np.where(inner, X + d[0], X)
    Y = np.where(inner, Y + d[1], Y)
    return np.stack([X.ravel(), Y.ravel()], axis=1)


def jittered_delaunay(n, jitter=0.3, seed=2, periodic=True, bc_code=401):
    """Delaunay triangulation of the jittered grid (scipy / Qhull).

    periodic: bottom 701 <-> top 702 (connection (0, 1)), left 711 <->
    right 712 (connection (1, 0)); otherwise every side gets bc_code.
    Returns (Raw, [BC])."""
    from scipy.spatial import Delaunay
    xy = jittered_points(n, jitter, seed)
    tri = Delaunay(xy, qhull_options="Qbb Qc Qz Q12").simplices.astype(np.int64)
    # drop degenerate slivers Qhull may emit on the collinear hull
    p = xy[tri]
    a2 = (p[:, 1, 0] - p[:, 0, 0]) * (p[:, 2, 1] - p[:, 0, 1]) - \
        (p[:, 2, 0] - p[:, 0, 0]) * (p[:, 1, 1] - p[:, 0, 1])
    tri = tri[np.abs(a2) > 1e-14 / (n * n)]
    tri4 = np.concatenate([tri, np.ones((len(tri), 1), np.int64)], axis=1)
    idx = np.arange(n * n).reshape(n, n)   # idx[j, i]
    bottom = np.stack([idx[0, :-1], idx[0, 1:]], 1)
    top = np.stack([idx[-1, :-1], idx[-1, 1:]], 1)
    left = np.stack([idx[:-1, 0], idx[1:, 0]], 1)
    right = np.stack([idx[:-1, -1], idx[1:, -1]], 1)
    if periodic:
        codes = (701, 702, 711, 712)
        bcs = [BC(701, connection=(0.0, 1.0)), BC(702),
               BC(711, connection=(1.0, 0.0)), BC(712)]
    else:
        codes = (bc_code,) * 4
        bcs = [BC(bc_code)]
    edges = []
    for e, c in zip((bottom, top, left, right), codes):
        edges.append(np.concatenate([e, np.full((len(e), 1), c)], 1))
    return Raw(xy, tri4, None, np.concatenate(edges)), bcs


def smooth_state(bary, seed=2, gamma=1.4, amp=0.1, mach=0.5):
    """Periodic smooth primitive state: rho, p = 1 + amp * sum of 4 sine
    modes with integer wave numbers in 1..4 (new phases for p), velocity of
    the given Mach number (relative to rho = p = 1) at a random angle."""
    rng = np.random.Generator(np.random.PCG64(seed + 1000))
    ab = rng.integers(1, 5, size=(4, 2))
    ph_r = rng.uniform(0, 2 * np.pi, 4)
    ph_p = rng.uniform(0, 2 * np.pi, 4)
    theta = rng.uniform(0, 2 * np.pi)
    x, y = bary[:, 0], bary[:, 1]
    sr = np.zeros_like(x)
    sp = np.zeros_like(x)
    for k in range(4):
        arg = 2 * np.pi * (ab[k, 0] * x + ab[k, 1] * y)
        sr += np.sin(arg + ph_r[k])
        sp += np.sin(arg + ph_p[k])
    u = mach * np.sqrt(gamma)
    prim = np.empty((len(x), 4))
    prim[:, 0] = 1.0 + amp * sr
    prim[:, 1] = u * np.cos(theta)
    prim[:, 2] = u * np.sin(theta)
    prim[:, 3] = 1.0 + amp * sp
    return prim


def naca4_surface(m, p, t, npts):
    """Closed-trailing-edge NACA 4-digit surface, npts points clockwise from
    the trailing edge (upper surface first), cosine spacing."""
    k = npts // 2
    beta = np.linspace(0.0, np.pi, k + 1)
    xc = 0.5 * (1.0 - np.cos(beta))          # 0 .. 1
    yt = 5 * t * (0.2969 * np.sqrt(xc) - 0.1260 * xc - 0.3516 * xc ** 2 +
                  0.2843 * xc ** 3 - 0.1036 * xc ** 4)
    yc = np.where(xc < p, m / p ** 2 * (2 * p * xc - xc ** 2),
                  m / (1 - p) ** 2 * ((1 - 2 * p) + 2 * p * xc - xc ** 2))
    dyc = np.where(xc < p, 2 * m / p ** 2 * (p - xc),
                   2 * m / (1 - p) ** 2 * (p - xc))
    th = np.arctan(dyc)
    xu, yu = xc - yt * np.sin(th), yc + yt * np.cos(th)
    xl, yl = xc + yt * np.sin(th), yc - yt * np.cos(th)
    # TE -> LE along the upper side, LE -> TE along the lower side
    x = np.concatenate([xu[::-1], xl[1:-1]])
    y = np.concatenate([yu[::-1], yl[1:-1]])
    return np.stack([x, y], 1)[:npts]


def naca_ogrid(ni=512, nj=256, m=0.02, p=0.4, t=0.12, radius=20.0,
               wall_spacing=1e-4, wall_bc=101, far_bc=501):
    """O-grid of ni x nj quads: ni points around the section (periodic in
    the wrap direction), nj cells radially with geometric stretching from
    wall_spacing out to a circle of the given radius (chords)."""
    surf = naca4_surface(m, p, t, ni)
    c = np.array([0.5, 0.0])
    ang = np.arctan2(surf[:, 1] - c[1], surf[:, 0] - c[0])
    outer = c + radius * np.stack([np.cos(ang), np.sin(ang)], 1)
    # geometric stretching ratio q with sum_{k<nj} ds q^k = radius
    def total(q):
        return wall_spacing * (q ** nj - 1) / (q - 1)
    lo, hi = 1.0 + 1e-9, 2.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if total(mid) < radius:
            lo = mid
        else:
            hi = mid
    q = 0.5 * (lo + hi)
    s = np.concatenate([[0.0], np.cumsum(wall_spacing * q ** np.arange(nj))])
    s = s / s[-1]
    # straight-ish lines blended from surface normal to the radial direction
    xy = (surf[None, :, :] * (1 - s[:, None, None]) +
          outer[None, :, :] * s[:, None, None]).reshape(-1, 2)
    idx = np.arange((nj + 1) * ni).reshape(nj + 1, ni)
    quads = []
    for j in range(nj):
        a = idx[j]
        b = np.roll(idx[j], -1)
        cc = np.roll(idx[j + 1], -1)
        d = idx[j + 1]
        quads.append(np.stack([a, b, cc, d, np.ones(ni, np.int64)], 1))
    quad = np.concatenate(quads)
    wall = np.stack([idx[0], np.roll(idx[0], -1),
                     np.full(ni, wall_bc)], 1)
    far = np.stack([idx[nj], np.roll(idx[nj], -1),
                    np.full(ni, far_bc)], 1)
    return Raw(xy, None, quad, np.concatenate([wall, far]))


def channel_bump(ni=300, nj=100, length=3.0, height=1.0, bump=0.1,
                 wall_bc=101, in_bc=501, out_bc=801):
    """Channel [0, length] x [0, height] with a circular-arc bump of the
    given thickness ratio on the lower wall between x = 1 and 2."""
    x = np.linspace(0.0, length, ni + 1)
    yb = np.zeros_like(x)
    on = (x > 1.0) & (x < 2.0)
    R = (0.25 + bump ** 2) / (2 * bump)
    yb[on] = np.sqrt(np.maximum(R ** 2 - (x[on] - 1.5) ** 2, 0.0)) - (R - bump)
    eta = np.linspace(0.0, 1.0, nj + 1)
    X = np.repeat(x[None, :], nj + 1, 0)
    Y = yb[None, :] + eta[:, None] * (height - yb[None, :])
    xy = np.stack([X.ravel(), Y.ravel()], 1)
    idx = np.arange((nj + 1) * (ni + 1)).reshape(nj + 1, ni + 1)
    a = idx[:-1, :-1].ravel()
    b = idx[:-1, 1:].ravel()
    cc = idx[1:, 1:].ravel()
    d = idx[1:, :-1].ravel()
    quad = np.stack([a, b, cc, d, np.ones_like(a)], 1)
    bot = np.stack([idx[0, :-1], idx[0, 1:], np.full(ni, wall_bc)], 1)
    top = np.stack([idx[-1, :-1], idx[-1, 1:], np.full(ni, wall_bc)], 1)
    left = np.stack([idx[:-1, 0], idx[1:, 0], np.full(nj, in_bc)], 1)
    right = np.stack([idx[:-1, -1], idx[1:, -1], np.full(nj, out_bc)], 1)
    return Raw(xy, None, quad, np.concatenate([bot, top, left, right]))
