"""GPU parity of the analysis consumers and the source term (src/analyze.c,
src/source.c) against the oracle.

* boundary side states of the last operator evaluation: bit for bit, every
  operator implementation, 1st and 2nd order, after fvTimeDerivative and
  after an RK step (the last stage's input state, as side->pVar holds);
* CL, CD and CP on an airfoil O-grid: bit for bit (the sums run in the
  reference's list order), after an explicit step; after an implicit step
  to the tolerance of the Newton solve;
* calcErrors: L1 and L2 to 1e-12 relative (block sums in a fixed order
  instead of the element order), Linf to 1e-13 (the device's sine and the
  C library's may differ by an ulp); in FP32-flux mode to 2e-6 absolute
  (the state is held in FP32);
* source term: dU/dt to 1e-13 relative of the oracle (device sine and
  cosine against the C library's), Euler and Navier-Stokes, triangles and
  quadrilaterals, every operator implementation, and RK steps (the source
  at each stage's time);
* record points: the element states bit for bit.
"""
import numpy as np
import pytest

from ccfd_amd import gpu as G
from ccfd_amd import mesh as M
from oracle import bindings as B
from tests import cases

pytestmark = pytest.mark.gpu

VARIANTS = [G.KV_SPLIT, G.KV_FUSED, G.KV_FUSED_FACE]
VIDS = ["split", "fused", "fused_face"]


def _wing_mesh(ni=96, nj=24):
    raw = M.naca_ogrid(ni=ni, nj=nj, wall_spacing=2e-3, radius=8.0,
                       wall_bc=101, far_bc=501)
    bc = M.BC(501, rho=1.0, mach=0.5, alpha=3.0, pressure=1.0)
    return M.Mesh(raw, [M.BC(101), bc])


def _wing_state(mesh, seed=3):
    gam = 1.4
    v = 0.5 * np.sqrt(gam)
    ref = np.array([1.0, v * np.cos(np.radians(3.0)),
                    v * np.sin(np.radians(3.0)), 1.0])
    rng = np.random.default_rng(seed)
    prim = np.tile(ref, (mesh.n_elems, 1))
    prim[:, 0] *= 1.0 + 0.05 * rng.standard_normal(mesh.n_elems)
    prim[:, 3] *= 1.0 + 0.05 * rng.standard_normal(mesh.n_elems)
    return prim, ref


def _pair(mesh, order, variant=G.KV_AUTO, flux=5, lim=2, ns=0, mu=0.0,
          nrk=1, torder=1, precision=G.FP64):
    prm = G.params(flux=flux, order=order, limiter=lim, venk_k=5.0,
                   navierstokes=ns, mu=mu, nrk=nrk, torder=torder,
                   precision=precision)
    cfg = B.make_cfg(flux=flux, order=order, limiter=lim, venk_k=5.0,
                     navierstokes=ns, mu=mu, nrk=nrk, torder=torder)
    gs = G.Solver(mesh, prm)
    gs.set_variant(variant)
    return gs, B.Oracle(mesh, cfg)


@pytest.mark.parametrize("order,variant", [(1, G.KV_AUTO)] +
                         [(2, v) for v in VARIANTS],
                         ids=["order1"] + VIDS)
def test_boundary_side_states_bitwise(order, variant):
    mesh, _ = cases.mixed_bc_mesh()
    gs, orc = _pair(mesh, order, variant)
    prim = cases.random_state(mesh.n_elems, seed=4, mach=0.6)
    gs.set_prim(prim)
    orc.set_prim(prim)
    gs.fvtd(0.25)
    o = orc.fvtd(0.25, full=True)
    got = gs.bc_sides()
    want = o["side"][mesh.ghost_slot]
    print("bc sides", order, variant, np.abs(got - want).max(), flush=True)
    assert np.array_equal(got, want)
    gs.close()


@pytest.mark.parametrize("variant", VARIANTS, ids=VIDS)
def test_wing_coefficients_bitwise_after_rk_steps(variant):
    mesh = _wing_mesh()
    prim, ref = _wing_state(mesh)
    gs, orc = _pair(mesh, 2, variant, flux=2, nrk=3, torder=2)
    gs.set_prim(prim)
    orc.set_prim(prim)
    nP, nS, xyphi = gs.wing_init(101, 0.8)
    slots, oP, oS = orc.wing_lists(101)
    assert (nP, nS) == (oP, oS) and nP + nS == 96
    t = 0.0
    for k in range(3):
        dt, _, _ = orc.calc_dt(t, 1e300, 1e300)
        orc.explicit_step(t, dt)
        gs.explicit_step(t, dt)
        t += dt
        cl, cd, cp = gs.wing_coef(ref, 3.0, nP + nS)
        ocl, ocd, ocp, oxy = orc.wing_coef(slots, oP, oS, 0.8, ref, 3.0)
        print("step", k, "cl", cl, ocl, "cd", cd, ocd, flush=True)
        assert cl == ocl and cd == ocd
        assert np.array_equal(cp, ocp)
    assert np.array_equal(xyphi, oxy)
    gs.close()


def test_wing_coefficients_after_an_implicit_step():
    mesh = _wing_mesh(64, 16)
    prim, ref = _wing_state(mesh, seed=5)
    gs, orc = _pair(mesh, 1, flux=2)
    gs.set_prim(prim)
    orc.set_prim(prim)
    nP, nS, _ = gs.wing_init(101)
    slots, oP, oS = orc.wing_lists(101)
    dt, _, _ = orc.calc_dt(0.0, 1e300, 1e300)
    ig = G.ImplicitCfg(20, 20, 1, 1e-8, 1e-4, 0.9)
    io = B.ImplicitCfg(20, 20, 1, 1e-8, 1e-4, 0.9)
    gs.implicit_step(ig, 0.0, 4 * dt)
    orc.implicit_step(io, 0.0, 4 * dt)
    cl, cd, cp = gs.wing_coef(ref, 3.0, nP + nS)
    ocl, ocd, ocp, _ = orc.wing_coef(slots, oP, oS, 1.0, ref, 3.0)
    print("implicit cl", cl, ocl, "cd", cd, ocd, flush=True)
    assert abs(cl - ocl) < 1e-8 and abs(cd - ocd) < 1e-8
    assert np.abs(cp - ocp).max() < 1e-7
    gs.close()


def _sine_mesh(kind):
    if kind == "tri":
        raw, _ = M.jittered_delaunay(20, seed=7, periodic=False, bc_code=601)
        return M.Mesh(raw, [M.BC(601, exact_func=3)])
    raw = M.cartesian(18, 12, (0.0, 0.0), (2.0, 1.5), (601, 601, 601, 601))
    return M.Mesh(raw, [M.BC(601, exact_func=3)])


@pytest.mark.parametrize("kind,order,precision",
                         [("tri", 2, G.FP64), ("quad", 1, G.FP64),
                          ("quad", 2, G.FP64), ("tri", 2, G.FP32_FLUX)],
                         ids=["tri_o2", "quad_o1", "quad_o2", "tri_o2_fp32"])
def test_error_norms_match_the_oracle(kind, order, precision):
    mesh = _sine_mesh(kind)
    gs, orc = _pair(mesh, order, flux=5, lim=1, precision=precision)
    prim = np.array([B.exact_func(orc.cfg, mesh, 3, x, 0.0)
                     for x in mesh.bary])
    prim[:, 0] += 0.01 * np.sin(7 * mesh.bary[:, 1])
    gs.set_prim(prim)
    orc.set_prim(prim)
    for t in (0.0, 0.13):
        L1, L2, Li = gs.calc_errors(3, t)
        o1, o2, oi = orc.calc_errors(3, t)
        print(kind, order, t, L1, o1, flush=True)
        if precision == G.FP64:
            np.testing.assert_allclose(L1, o1, rtol=1e-12, atol=1e-300)
            np.testing.assert_allclose(L2, o2, rtol=1e-12, atol=1e-300)
            np.testing.assert_allclose(Li, oi, rtol=1e-13, atol=1e-15)
        else:
            # the FP32 state itself is rounded to ~5e-7 (values near 3)
            for a, b in ((L1, o1), (L2, o2), (Li, oi)):
                np.testing.assert_allclose(a, b, rtol=0, atol=2e-6)
    # the Riemann problem (exact function 5) on the quadrilaterals
    if kind == "quad" and precision == G.FP64:
        L1, L2, Li = gs.calc_errors(5, 0.07)
        o1, o2, oi = orc.calc_errors(5, 0.07)
        np.testing.assert_allclose(L1, o1, rtol=1e-12)
        np.testing.assert_allclose(Li, oi, rtol=1e-13)
    gs.close()


@pytest.mark.parametrize("kind,ns,order,variant",
                         [("tri", 0, 2, G.KV_SPLIT), ("tri", 0, 2, G.KV_FUSED),
                          ("tri", 0, 2, G.KV_FUSED_FACE),
                          ("quad", 0, 1, G.KV_AUTO), ("quad", 1, 2, G.KV_AUTO)],
                         ids=["tri_split", "tri_fused", "tri_face",
                              "quad_o1", "quad_ns"])
def test_source_term_matches_the_oracle(kind, ns, order, variant):
    if kind == "tri":
        raw, bcs = M.jittered_delaunay(16, seed=3)
        mesh = M.Mesh(raw, bcs)
    else:
        raw = M.cartesian(12, 9, (0.0, 0.0), (1.0, 0.8),
                          (101, 401, 101, 401))
        mesh = M.Mesh(raw, [M.BC(101), M.BC(401)])
    mu = 0.01 if ns else 0.0
    prm = G.params(flux=2, order=order, limiter=2, venk_k=5.0,
                   navierstokes=ns, mu=mu, nrk=4, torder=2)
    cfg = B.make_cfg(flux=2, order=order, limiter=2, venk_k=5.0,
                     navierstokes=ns, mu=mu, nrk=4, torder=2, source=1)
    gs = G.Solver(mesh, prm)
    gs.set_variant(variant)
    gs.set_source(1)
    orc = B.Oracle(mesh, cfg)
    prim = cases.random_state(mesh.n_elems, seed=6, mach=0.4)
    gs.set_prim(prim)
    orc.set_prim(prim)
    gs.fvtd(0.37)
    want = orc.fvtd(0.37)
    err = cases.rel_err(gs.ut(), want)
    print("source fvtd rel err", err, flush=True)
    assert err < 1e-13
    t = 0.0
    for _ in range(3):
        dt, _, _ = orc.calc_dt(t, 1e300, 1e300)
        gs.explicit_step(t, dt)
        orc.explicit_step(t, dt)
        t += dt
    err = cases.rel_err(gs.prim(), orc.prim())
    print("source RK4 rel err", err, flush=True)
    assert err < 1e-13
    # switching the source off gives the plain operator again, bit for bit
    gs.set_source(0)
    gs.fvtd(0.37)
    orc0 = B.Oracle(mesh, B.make_cfg(flux=2, order=order, limiter=2,
                                     venk_k=5.0, navierstokes=ns, mu=mu))
    orc0.set_prim(gs.prim())
    assert np.array_equal(gs.ut(), orc0.fvtd(0.37))
    gs.close()


def test_record_points_bitwise():
    mesh, _ = cases.mixed_bc_mesh()
    gs, _ = _pair(mesh, 2, flux=5, lim=1)
    prim = cases.random_state(mesh.n_elems, seed=8)
    gs.set_prim(prim)
    idx = np.array([0, 5, mesh.n_elems - 1, 17, 5])
    assert np.array_equal(gs.prim_at(idx), prim[idx])
    gs.explicit_step(0.0, 1e-4)
    assert np.array_equal(gs.prim_at(idx[:2]), gs.prim()[idx[:2]])
    with pytest.raises(G.GpuError):
        gs.prim_at([mesh.n_elems])
    gs.close()
