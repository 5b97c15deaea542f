"""Request-time mesh preprocessing on the device (ccfd_batch_add_raw).

An instance admitted from raw connectivity - sides paired, geometry,
least-squares weights, element order and block halos all built on the GPU -
must behave bit for bit as the instance admitted from the host builder's
image (ccfd_mesh_build + build_soa, i.e. src/mesh.c's construction): the
same time step, residuals and states after several steps, in FP64 and in
FP32-flux arenas, for triangle, quadrilateral and mixed meshes, next to other
instances, periodic meshes included. Malformed meshes are refused with the
host builder's messages."""
import math

import numpy as np
import pytest

from ccfd_amd import batch as BT
from ccfd_amd import gpu as G
from ccfd_amd import mesh as M
from tests import cases

pytestmark = pytest.mark.gpu

MB = 1 << 20


def _bump_bcs(mach=0.6):
    return [M.BC(101), M.BC(501, rho=1.0, mach=mach, alpha=0.0, pressure=1.0),
            M.BC(801, pressure=1.0)]


def _case(name):
    """(raw, bcs, max sides, local time steps)"""
    if name == "delaunay":
        raw, bcs = M.jittered_delaunay(40, 0.3, 7, periodic=False, bc_code=401)
        return raw, bcs, 3, False
    if name == "periodic":
        # connectPeriodicBC on the device: both directions periodic
        raw, bcs = M.jittered_delaunay(48, 0.3, 9, periodic=True)
        return raw, bcs, 3, False
    if name == "naca":
        raw = M.naca_ogrid(96, 48, wall_bc=101, far_bc=501)
        bcs = [M.BC(101), M.BC(501, rho=1.0, mach=0.7, alpha=2.0, pressure=1.0)]
        return raw, bcs, 4, True
    if name == "bump":
        raw = M.channel_bump(90, 30, wall_bc=101, in_bc=501, out_bc=801)
        return raw, _bump_bcs(), 4, True
    if name == "mixed":
        raw = M.channel_bump(60, 20, wall_bc=101, in_bc=501, out_bc=801)
        return cases.split_quads(raw, seed=3), _bump_bcs(0.8), 4, False
    if name == "large":
        # 49,152 quads: 192 blocks and their halo lists
        raw = M.naca_ogrid(256, 192, wall_bc=101, far_bc=501)
        bcs = [M.BC(101), M.BC(501, rho=1.0, mach=0.5, alpha=4.0, pressure=1.0)]
        return raw, bcs, 4, True
    raise ValueError(name)


def _state(mesh, seed):
    """free stream with a seeded smooth perturbation, so every flux and
    limiter branch sees varied states"""
    rng = np.random.Generator(np.random.PCG64(seed))
    x, y = mesh.bary[:, 0], mesh.bary[:, 1]
    k = rng.uniform(1.0, 3.0, size=4)
    ph = rng.uniform(0.0, 2 * math.pi, size=4)
    d = 0.05 * np.sin(k[0] * x + ph[0]) * np.cos(k[1] * y + ph[1])
    c = math.sqrt(1.4)
    u = 0.6 * c + 0.1 * np.sin(k[2] * y + ph[2])
    v = 0.1 * np.cos(k[3] * x + ph[3])
    return np.stack([1.0 + d, u, v, 1.0 + 0.5 * d], axis=1)


@pytest.mark.parametrize("prec", [G.FP64, G.FP32_FLUX], ids=["fp64", "fp32"])
@pytest.mark.parametrize("name", ["delaunay", "periodic", "naca", "bump", "mixed",
                                  "large"])
def test_device_preprocessing_equals_the_host_image(prec, name):
    raw, bcs, sides, lts = _case(name)
    host = M.Mesh(raw, bcs)
    prim = _state(host, 11)
    sch = G.params(flux=9, order=2, limiter=2, venk_k=5.0, nrk=3, torder=2,
                   precision=prec)
    b = BT.Batch(sch, max_sides=sides, local_dt=lts, hbm_bytes=256 * MB)
    b.set_variant(G.KV_FUSED_FACE)
    # a neighbour instance before and after, so the image lands at offsets
    other = cases.delaunay_mesh(12, seed=2, periodic=False, bc_code=401)
    k0 = b.add(other, _state(other, 3), cfl=0.5)
    kh = b.add(host, prim, cfl=0.8)
    print("case %s: %d elements, %d ghosts, device build" % (
        name, host.n_elems, host.n_ghosts), flush=True)
    kd = b.add_raw(M.RawMesh(raw, bcs), prim, cfl=0.8)
    k1 = b.add(other, _state(other, 4), cfl=0.5)
    b.step(5)
    b.sync()
    ih, idv = b.info(kh), b.info(kd)
    print("host res %s, device res %s" % (ih["res"], idv["res"]))
    assert idv["n"] == ih["n"] == host.n_elems
    assert idv["steps"] == ih["steps"] == 5
    assert idv["t"] == ih["t"] and idv["dt"] == ih["dt"]
    assert np.array_equal(idv["res0"], ih["res0"])
    assert np.array_equal(idv["res"], ih["res"])
    assert np.array_equal(b.prim(kd), b.prim(kh))
    assert np.array_equal(b.cons(kd), b.cons(kh))
    for k in (k0, k1):
        assert b.info(k)["steps"] == 5
    b.close()


def test_device_preprocessing_refusals():
    sch = G.params(flux=9, order=2, limiter=2, venk_k=5.0, nrk=3, torder=2)
    b = BT.Batch(sch, max_sides=4, hbm_bytes=64 * MB)
    raw = M.cartesian(6, 4, side_bc=(101, 401, 101, 401))
    bcs = [M.BC(101), M.BC(401)]
    prim = np.tile([1.0, 0.5, 0.0, 1.0], (len(raw.quad), 1))
    k = b.add_raw(M.RawMesh(raw, bcs), prim)      # the good mesh is admitted
    assert b.info(k)["n"] == len(raw.quad)

    def refused(r, bc, match):
        with pytest.raises(BT.GpuError, match=match):
            b.add_raw(M.RawMesh(r, bc), prim)

    # two boundary edges fewer: the count stays even, two sides go unmatched
    refused(M.Raw(raw.xy, quad=raw.quad, bcedge=raw.bcedge[2:]), bcs,
            "unmatched side")
    # a boundary code the parameter list does not define
    refused(raw, [M.BC(101)], "not defined")
    bad = raw.quad.copy()
    bad[0, 0] = len(raw.xy)
    refused(M.Raw(raw.xy, quad=bad, bcedge=raw.bcedge), bcs, "references node")
    refused(M.Raw(raw.xy, quad=raw.quad, bcedge=raw.bcedge[1:]), bcs,
            "open mesh")
    # a periodic pair whose connection vector misses its partner
    per, pbcs = M.jittered_delaunay(8, 0.3, 1, periodic=True)
    pbcs = [M.BC(c.code, connection=(0.0, 1.5)) if c.code == 701 else c
            for c in pbcs]
    b3 = BT.Batch(sch, max_sides=3, hbm_bytes=32 * MB)
    with pytest.raises(BT.GpuError, match="no connection"):
        b3.add_raw(M.RawMesh(per, pbcs),
                   np.tile([1.0, 0.5, 0.0, 1.0], (len(per.tri), 1)))
    with pytest.raises(ValueError, match="no connection"):
        M.Mesh(per, pbcs)
    # the refusals left the arena as it was
    b.step(2)
    b.sync()
    assert b.info(k)["steps"] == 2
    used, _, n_inst = b.usage()
    assert n_inst == 1 and used == 256     # one block
    b.close()
    b3.close()
