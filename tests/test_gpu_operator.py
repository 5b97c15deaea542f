"""GPU parity of the spatial operator (fvTimeDerivative) against the oracle.

FP64: bit-identical (the drop-in bar is 1e-12 relative; the HIP kernels
keep the reference's operation order, no FMA contraction, IEEE div/sqrt).
FP32-flux: within 1e-4 relative of the FP64 oracle.
"""
import numpy as np
import pytest

from ccfd_amd import gpu as G
from ccfd_amd import mesh as M
from oracle import bindings as B
from tests import cases
from tests import golden_io

pytestmark = pytest.mark.gpu


def _pair(mesh, flux, order, lim, k=5.0, ns=0, mu=0.0, precision=G.FP64,
          variant=G.KV_AUTO):
    prm = G.params(flux=flux, order=order, limiter=lim, venk_k=k,
                   navierstokes=ns, mu=mu, precision=precision)
    cfg = B.make_cfg(flux=flux, order=order, limiter=lim, venk_k=k,
                     navierstokes=ns, mu=mu)
    gs = G.Solver(mesh, prm)
    gs.set_variant(variant)
    if variant != G.KV_AUTO:
        assert gs.variant() == variant
    elif order == 1 or ns:
        assert gs.variant() == G.KV_SPLIT
    return gs, B.Oracle(mesh, cfg)


# the 2nd-order Euler operator is checked in both implementations
VARIANTS = pytest.mark.parametrize(
    "variant", [G.KV_SPLIT, G.KV_FUSED, G.KV_FUSED_FACE],
    ids=["split", "fused", "fused_face"])
# (order, limiter, operator implementation): the fused operators are 2nd
# order only, so 1st order runs the split pair alone
ORDER_VARIANTS = pytest.mark.parametrize(
    "order,lim,variant",
    [(1, 1, G.KV_SPLIT)] + [(2, lim, v) for lim in (1, 2)
                            for v in (G.KV_SPLIT, G.KV_FUSED, G.KV_FUSED_FACE)],
    ids=["o1-split"] + ["o2-l%d-%s" % (lim, n) for lim in (1, 2)
                        for n in ("split", "fused", "fused_face")])


def _check_bitwise(gs, orc, prim, time=0.0, order=2):
    gs.set_prim(prim)
    orc.set_prim(prim)
    gs.fvtd(time)
    o = orc.fvtd(time, full=True)
    ut = gs.ut()
    print("max |du/dt diff| =", np.abs(ut - o["u_t"]).max(), flush=True)
    assert np.array_equal(ut, o["u_t"])
    ux, uy, sp = gs.recon()
    if order == 2:
        assert np.array_equal(ux, o["u_x"])
        assert np.array_equal(uy, o["u_y"])
    assert np.array_equal(sp, o["side"])


@pytest.mark.parametrize("code", list(range(1, 12)))
def test_riemann_batch_bitwise(code):
    rng = np.random.default_rng(code)
    n = 4096
    def states():
        s = np.empty((n, 4))
        s[:, 0] = np.exp(rng.uniform(-3, 3, n))
        s[:, 3] = np.exp(rng.uniform(-3, 3, n))
        c = np.sqrt(1.4 * s[:, 3] / s[:, 0])
        s[:, 1] = c * rng.uniform(-3, 3, n)
        s[:, 2] = rng.normal(size=n)
        return s
    L, R = states(), states()
    Fg = G.riemann_batch(code, L, R)
    Fo = B.riemann(B.make_cfg(), code, L, R)
    if code == 1:
        # Godunov iterates with pow(): device libm may differ in the last ulp
        np.testing.assert_allclose(Fg, Fo, rtol=1e-10, atol=1e-12)
    else:
        assert np.array_equal(Fg, Fo, equal_nan=True)


@pytest.mark.parametrize("flux", [2, 3, 5, 9])
@pytest.mark.parametrize("lim", [1, 2])
@VARIANTS
def test_fvtd_sod_bitwise(flux, lim, variant):
    mesh = cases.sod_mesh(100, 1)
    gs, orc = _pair(mesh, flux, 2, lim, k=10.0, variant=variant)
    prim = cases.random_state(mesh.n_elems, seed=flux)
    _check_bitwise(gs, orc, prim, order=2)
    gs.close()


@pytest.mark.parametrize("flux", [2, 3, 5, 9])
def test_fvtd_sod_first_order_bitwise(flux):
    mesh = cases.sod_mesh(100, 1)
    gs, orc = _pair(mesh, flux, 1, 1, k=10.0)
    prim = cases.random_state(mesh.n_elems, seed=flux)
    _check_bitwise(gs, orc, prim, order=1)
    gs.close()


@pytest.mark.parametrize("flux", list(range(1, 12)))
@VARIANTS
def test_fvtd_periodic_delaunay_bitwise(flux, variant):
    mesh = cases.delaunay_mesh(40, seed=3, periodic=True)
    gs, orc = _pair(mesh, flux, 2, 2, k=5.0, variant=variant)
    prim = M.smooth_state(mesh.bary, seed=2)
    if flux == 1:
        # Godunov iterates with pow(): device libm may differ in the last ulp
        gs.set_prim(prim)
        orc.set_prim(prim)
        gs.fvtd(0.0)
        np.testing.assert_allclose(gs.ut(), orc.fvtd(0.0), rtol=1e-10,
                                   atol=1e-12)
    else:
        _check_bitwise(gs, orc, prim)
    gs.close()


@ORDER_VARIANTS
def test_fvtd_all_bc_types_bitwise(order, lim, variant):
    mesh, _ = cases.mixed_bc_mesh()
    gs, orc = _pair(mesh, 5, order, lim, k=5.0, variant=variant)
    prim = cases.random_state(mesh.n_elems, seed=11, mach=0.5)
    _check_bitwise(gs, orc, prim, order=order)
    gs.close()


@ORDER_VARIANTS
def test_fvtd_mixed_triangles_and_quads_bitwise(order, lim, variant):
    """elements with 3 and 4 sides in one mesh: the absent fourth slot of
    every triangle is skipped by every kernel"""
    mesh, _ = cases.mixed_mesh()
    assert len(mesh.raw.tri) > 100 and len(mesh.raw.quad) > 100
    gs, orc = _pair(mesh, 9, order, lim, k=5.0, variant=variant)
    prim = cases.random_state(mesh.n_elems, seed=17, mach=0.5)
    _check_bitwise(gs, orc, prim, order=order)
    gs.close()


GOLDEN = golden_io.operator_cases()


@pytest.mark.parametrize("name", sorted(GOLDEN))
def test_fvtd_matches_reference_outputs(name):
    """the device operator against the reference's own fvTimeDerivative
    outputs in the committed fixtures, with no oracle in between"""
    mesh, _, prim, c = golden_io.build(GOLDEN[name])
    prm = G.params(flux=c["flux"], order=c["order"], limiter=c["limiter"],
                   venk_k=c["venk_k"], navierstokes=c["ns"], mu=c["mu"],
                   left=c.get("left", (1, 0, 0, 1)),
                   right=c.get("right", (0.125, 0, 0, 0.1)),
                   x0=c.get("x0", 0.5))
    gs = G.Solver(mesh, prm)
    gs.set_prim(prim)
    gs.fvtd(c["time"])
    g = GOLDEN[name]
    ut = gs.ut()
    print("%s: %d elements, max |du/dt diff| = %.3e" % (
        name, mesh.n_elems, np.abs(ut - g["u_t"]).max()), flush=True)
    if c["flux"] == 1:
        # Godunov iterates with pow(): device libm may differ in the last ulp
        np.testing.assert_allclose(ut, g["u_t"], rtol=1e-10, atol=1e-12)
    else:
        assert np.array_equal(ut, g["u_t"])
        ux, uy, sp = gs.recon()
        if c["order"] == 2:
            assert np.array_equal(ux, g["u_x"])
            assert np.array_equal(uy, g["u_y"])
        assert np.array_equal(sp, g["side"])
    gs.close()


@pytest.mark.parametrize("order,variant", [(1, G.KV_AUTO), (2, G.KV_SPLIT),
                                           (2, G.KV_FUSED),
                                           (2, G.KV_FUSED_FACE)],
                         ids=["order1", "split", "fused", "fused_face"])
def test_exact_function_boundaries_on_every_path(order, variant):
    """exact-function boundaries (type 6) are set by their own kernel after
    the common boundary kernels: every operator implementation and both
    orders stay bit-identical to the oracle (mesh and settings of the
    reference-generated exact-boundary Sod fixture)"""
    g = GOLDEN["exactbc_sod_t01_o2_hll"]
    mesh, _, prim, c = golden_io.build(g)
    assert any(int(r[0]) // 100 == 6 for r in g["bcs"])
    kw = dict(left=c["left"], right=c["right"], x0=c["x0"])
    prm = G.params(flux=c["flux"], order=order, limiter=c["limiter"],
                   venk_k=c["venk_k"], **kw)
    cfg = B.make_cfg(flux=c["flux"], order=order, limiter=c["limiter"],
                     venk_k=c["venk_k"], **kw)
    gs = G.Solver(mesh, prm)
    gs.set_variant(variant)
    orc = B.Oracle(mesh, cfg)
    _check_bitwise(gs, orc, prim, time=c["time"], order=order)
    gs.close()


def test_fused_rejected_where_it_does_not_apply():
    raw = M.cartesian(8, 4, (0.0, 0.0), (1.0, 0.5), (201, 501, 101, 401))
    bcs = [M.BC(201), M.BC(501, rho=1.0, mach=0.3, alpha=0.0, pressure=1.0),
           M.BC(101), M.BC(401)]
    mesh = M.Mesh(raw, bcs)
    prm = G.params(flux=2, order=2, limiter=2, navierstokes=1, mu=0.01)
    gs = G.Solver(mesh, prm)
    assert gs.variant() == G.KV_SPLIT
    for v in (G.KV_FUSED, G.KV_FUSED_FACE):
        gs.set_variant(v)
        with pytest.raises(G.GpuError):
            gs.variant()
    gs.close()


@pytest.mark.parametrize("order", [1, 2])
def test_fvtd_navier_stokes_bitwise(order):
    raw = M.cartesian(20, 10, (0.0, 0.0), (1.0, 0.5), (201, 501, 101, 401))
    bcs = [M.BC(201), M.BC(501, rho=1.0, mach=0.3, alpha=0.0, pressure=1.0),
           M.BC(101), M.BC(401)]
    mesh = M.Mesh(raw, bcs)
    gs, orc = _pair(mesh, 2, order, 2, k=5.0, ns=1, mu=0.01)
    prim = cases.random_state(mesh.n_elems, seed=5, mach=0.3)
    _check_bitwise(gs, orc, prim, order=order)
    gs.close()


@pytest.mark.parametrize("flux", [2, 3, 5, 9])
@VARIANTS
def test_fvtd_fp32_flux_tolerance(flux, variant):
    mesh = cases.delaunay_mesh(40, seed=3, periodic=True)
    gs, orc = _pair(mesh, flux, 2, 2, k=5.0, precision=G.FP32_FLUX,
                    variant=variant)
    prim = M.smooth_state(mesh.bary, seed=2)
    gs.set_prim(prim)
    orc.set_prim(prim)
    gs.fvtd(0.0)
    err = cases.rel_err(gs.ut(), orc.fvtd(0.0))
    print("fp32 rel err", err)
    assert err < 1e-4
    gs.close()


@VARIANTS
def test_explicit_rk_steps_bitwise(variant):
    """Sod, HLL, 2nd order VK, 3-stage RK: ten steps of calcTimeStep +
    explicitTimeStepRK match the oracle bit for bit (state, dt, residual)."""
    mesh = cases.sod_mesh(100, 1)
    prm = G.params(flux=3, order=2, limiter=2, venk_k=10.0, ts1d=1, nrk=3,
                   torder=2)
    cfg = B.make_cfg(flux=3, order=2, limiter=2, venk_k=10.0, ts1d=1, nrk=3,
                     torder=2)
    gs, orc = G.Solver(mesh, prm), B.Oracle(mesh, cfg)
    gs.set_variant(variant)
    assert gs.variant() == variant
    prim = cases.sod_state(mesh)
    gs.set_prim(prim)
    orc.set_prim(prim)
    t = 0.0
    for _ in range(10):
        dtg, _ = gs.calc_dt(t, 0.25, 0.25)
        dto, _, _ = orc.calc_dt(t, 0.25, 0.25)
        assert dtg == dto
        rg = gs.explicit_step(t, dtg)
        ro = orc.explicit_step(t, dto)
        np.testing.assert_allclose(rg, ro, rtol=1e-13)
        t += dtg
    assert np.array_equal(gs.cons(), orc.cons())
    assert np.array_equal(gs.prim(), orc.prim())
    gs.close()


@VARIANTS
def test_advance_device_dt_matches_host_loop(variant):
    """ccfd_gpu_advance (dt kept on the device) equals the host-driven loop."""
    mesh = cases.delaunay_mesh(30, seed=4, periodic=True)
    prm = G.params(flux=5, order=2, limiter=2, venk_k=5.0, nrk=4, torder=2)
    a = G.Solver(mesh, prm)
    b = G.Solver(mesh, prm)
    b.set_variant(variant)
    prim = M.smooth_state(mesh.bary, seed=2)
    a.set_prim(prim)
    b.set_prim(prim)
    t = 0.0
    for _ in range(5):
        dt, _ = a.calc_dt(t, 1e9, 1e9)
        a.explicit_step(t, dt)
        t += dt
    t2, _ = b.advance(5, 0.0, 1e9, 1e9)
    assert t2 == t
    assert np.array_equal(a.cons(), b.cons())
    a.close()
    b.close()


@pytest.mark.parametrize("case", ["fp64_tile512", "fp32_face", "ns_split",
                                  "ts1d_quads"])
def test_device_time_steps_equal_the_host_loop(case):
    """ccfd_gpu_advance (calcTimeStep and the steps on the device, several
    steps per call) equals the host loop that calls calcTimeStep before
    every step, bit for bit: 512-element tiles, FP32 flux, the viscous
    limit (Navier-Stokes, split operator) and the 1-D time step on
    quadrilaterals"""
    kw = dict(flux=5, order=2, limiter=2, venk_k=5.0, nrk=4, torder=2)
    tile = 0
    if case in ("fp64_tile512", "fp32_face"):
        mesh = cases.delaunay_mesh(120, seed=5, periodic=True)
        prim = M.smooth_state(mesh.bary, seed=3)
        if case == "fp32_face":
            kw["precision"] = G.FP32_FLUX
        else:
            tile = 512
    elif case == "ns_split":
        raw = M.naca_ogrid(ni=64, nj=16, wall_spacing=4e-3, radius=6.0,
                           wall_bc=201, far_bc=501)
        mesh = M.Mesh(raw, [M.BC(201), M.BC(501, rho=1.0, mach=0.4, alpha=2.0,
                                              pressure=1.0)])
        prim = np.tile([1.0, 0.4 * np.sqrt(1.4), 0.0, 1.0], (mesh.n_elems, 1))
        kw.update(flux=2, navierstokes=1, mu=2e-3)
    else:
        raw = M.cartesian(40, 3, (0.0, 0.0), (1.0, 0.05), (101, 401, 101, 401))
        mesh = M.Mesh(raw, [M.BC(101), M.BC(401)])
        prim = cases.sod_state(mesh)
        kw.update(ts1d=1, nrk=3)
    a = G.Solver(mesh, G.params(**kw))
    b = G.Solver(mesh, G.params(**kw))
    if tile:
        b.set_variant(G.KV_FUSED_FACE)
        b.set_tile(tile)
        assert b.tile() == tile
    a.set_prim(prim)
    b.set_prim(prim)
    t = 0.0
    for _ in range(4):
        dt, _ = a.calc_dt(t, 1e9, 1e9)
        a.explicit_step(t, dt)
        t += dt
    t2, _ = b.advance(4, 0.0, 1e9, 1e9)
    print(case, t, t2, flush=True)
    assert t2 == t
    assert np.array_equal(a.cons(), b.cons())
    a.close()
    b.close()


@pytest.mark.parametrize("n,periodic", [(120, True), (300, True),
                                        (200, False)])
def test_fused_equals_split_large_delaunay(n, periodic):
    """Fused and split operators, and the face-once kernel with tiles of
    256 and 512 elements, agree bit for bit on meshes with many blocks
    (every block has a halo) through RK4 steps, residuals included;
    transmissive boundaries when not periodic."""
    mesh = cases.delaunay_mesh(n, seed=7, periodic=periodic)
    prm = G.params(flux=5, order=2, limiter=2, venk_k=5.0, nrk=4, torder=2)
    prim = M.smooth_state(mesh.bary, seed=9)
    out = []
    for v, tile in ((G.KV_SPLIT, 0), (G.KV_FUSED, 0), (G.KV_FUSED_FACE, 256),
                    (G.KV_FUSED_FACE, 512)):
        a = G.Solver(mesh, prm)
        a.set_variant(v)
        assert a.variant() == v
        if tile:
            a.set_tile(tile)
            assert a.tile() == tile
        a.set_prim(prim)
        t, r = a.advance(3, 0.0, 1e9, 1e9)
        out.append((t, r, a.cons(), a.ut()))
        a.close()
    for t, r, cons, ut in out[1:]:
        assert t == out[0][0]
        assert np.array_equal(r, out[0][1])
        assert np.array_equal(cons, out[0][2])
        assert np.array_equal(ut, out[0][3])


@pytest.fixture(scope="module")
def large_delaunay():
    """250,632 periodic jittered-Delaunay triangles (355 points per side),
    the smooth seeded state of the 4M-cell workload"""
    mesh = cases.delaunay_mesh(355, seed=2, periodic=True)
    return mesh, M.smooth_state(mesh.bary, seed=2)


@pytest.mark.parametrize("flux", [2, 3, 5, 9], ids=["roe", "hll", "hllc",
                                                     "ausmd"])
def test_large_delaunay_matches_the_oracle(large_delaunay, flux):
    """fvTimeDerivative (2nd order, VK K = 5) on 250k periodic triangles,
    the default (fused face-once) operator: dU/dt equal to the oracle's bit
    for bit, then two RK4 steps with device time steps and the residual of
    each step, equal to the oracle's steps"""
    mesh, prim = large_delaunay
    kw = dict(flux=flux, order=2, limiter=2, venk_k=5.0, nrk=4, torder=2)
    gs = G.Solver(mesh, G.params(**kw))
    assert gs.variant() == G.KV_FUSED_FACE
    orc = B.Oracle(mesh, B.make_cfg(**kw))
    gs.set_prim(prim)
    orc.set_prim(prim)
    gs.fvtd(0.0)
    assert np.array_equal(gs.ut(), orc.fvtd(0.0))
    t, to = 0.0, 0.0
    for _ in range(2):
        t, r = gs.advance(1, t, 1e9, 1e9)
        dto, _, _ = orc.calc_dt(to, 1e9, 1e9)
        ro = orc.explicit_step(to, dto)
        to += dto
        assert t == to
        assert np.allclose(r, ro, rtol=1e-12, atol=0.0)
    assert np.array_equal(gs.cons(), orc.cons())
    gs.close()


def test_autotune_keeps_the_state_and_picks_a_timed_candidate():
    """ccfd_gpu_autotune times every applicable implementation and tile on
    the resident state without changing it, keeps the fastest, and the
    operator then gives the same dU/dt as before"""
    mesh = cases.delaunay_mesh(150, seed=5, periodic=True)
    prim = M.smooth_state(mesh.bary, seed=4)
    gs = G.Solver(mesh, G.params(flux=5, order=2, limiter=2, venk_k=5.0))
    gs.set_prim(prim)
    gs.fvtd(0.0)
    ut0 = gs.ut()
    v, tile, ms = gs.autotune(3)
    print("autotune: variant %d tile %d %s" % (v, tile, ms))
    timed = {k: t for k, t in ms.items() if t is not None}
    assert {"split", "fused", "fused_face_256"} <= set(timed)
    assert all(t > 0 for t in timed.values())
    name = {G.KV_SPLIT: "split", G.KV_FUSED: "fused"}.get(
        v, "fused_face_%d" % tile)
    assert timed[name] <= min(timed.values()) / 0.98
    assert gs.variant() == v and gs.tile() == tile
    assert np.array_equal(gs.prim(), prim) or np.allclose(gs.prim(), prim,
                                                          rtol=1e-15)
    gs.fvtd(0.0)
    assert np.array_equal(gs.ut(), ut0)
    gs.close()
