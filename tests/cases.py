"""Shared test cases: meshes, states and comparison helpers."""
import numpy as np

from ccfd_amd import mesh as M

SOD_LEFT = (1.0, 0.0, 0.0, 1.0)
SOD_RIGHT = (0.125, 0.0, 0.0, 0.1)


def sod_mesh(ni=100, nj=1):
    """ccfd's make-check geometry: [0,1]x[0,0.01], slip walls 101 top and
    bottom, outflow 401 left and right (check/sod_*.ini)."""
    raw = M.cartesian(ni, nj, (0.0, 0.0), (1.0, 0.01), (101, 401, 101, 401))
    return M.Mesh(raw, [M.BC(101), M.BC(401)])


def sod_state(mesh, x0=0.5):
    """exactFunc 5 at t = 0 (src/exactFunction.c:131-140)"""
    x = mesh.bary[:, 0]
    prim = np.empty((mesh.n_elems, 4))
    left = x <= x0
    prim[left] = SOD_LEFT
    prim[~left] = SOD_RIGHT
    return prim


def random_state(n, seed=0, mach=0.8):
    rng = np.random.default_rng(seed)
    p = np.empty((n, 4))
    p[:, 0] = 1.0 + 0.6 * rng.random(n)
    p[:, 3] = 1.0 + 0.6 * rng.random(n)
    c = np.sqrt(1.4 * p[:, 3] / p[:, 0])
    ang = rng.uniform(0, 2 * np.pi, n)
    m = mach * rng.random(n)
    p[:, 1] = m * c * np.cos(ang)
    p[:, 2] = m * c * np.sin(ang)
    return p


def delaunay_mesh(n=40, seed=2, periodic=True, bc_code=401):
    raw, bcs = M.jittered_delaunay(n, 0.3, seed, periodic=periodic,
                                   bc_code=bc_code)
    return M.Mesh(raw, bcs)


def mixed_bc_mesh(ni=24, nj=12):
    """Cartesian mesh with a different BC on every side: slip wall bottom,
    pressure outlet right, inflow top... exercises every ghost-state type."""
    raw = M.cartesian(ni, nj, (0.0, 0.0), (2.0, 1.0), (101, 801, 301, 501))
    bcs = [M.BC(101), M.BC(801, pressure=0.9),
           M.BC(301, rho=1.1, mach=0.6, alpha=10.0, pressure=1.05),
           M.BC(501, rho=0.95, mach=0.4, alpha=-5.0, pressure=1.0)]
    return M.Mesh(raw, bcs), bcs


def rel_err(a, b):
    """max |a - b| / max(|b|, 1e-300) over the array, with a per-variable
    scale (the scale of dU/dt differs by variable)"""
    scale = np.maximum(np.abs(b).max(axis=0), 1e-300)
    return float((np.abs(a - b) / scale).max())


def split_quads(raw, seed=0, frac=0.5):
    """a mixed mesh: a seeded random subset of the quadrilaterals of `raw`
    split into two triangles along a diagonal (orientation kept)"""
    rng = np.random.Generator(np.random.PCG64(seed))
    q = raw.quad
    pick = rng.uniform(size=len(q)) < frac
    keep = q[~pick]
    sq = q[pick]
    alt = rng.uniform(size=len(sq)) < 0.5
    t1 = np.where(alt[:, None], sq[:, [0, 1, 2, 4]], sq[:, [0, 1, 3, 4]])
    t2 = np.where(alt[:, None], sq[:, [0, 2, 3, 4]], sq[:, [1, 2, 3, 4]])
    tri = np.concatenate([t1, t2])
    return M.Raw(raw.xy, tri, keep, raw.bcedge)


def mixed_mesh(seed=3):
    """the mixed-BC Cartesian mesh with about half its quadrilaterals split
    into triangles: elements with 3 and 4 sides side by side"""
    base, bcs = mixed_bc_mesh()
    return M.Mesh(split_quads(base.raw, seed=seed), bcs), bcs
