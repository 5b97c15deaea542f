"""Implicit path on the GPU (JFNK-GMRES + LU-SGS) against the oracle.

The oracle restates implicitTimeStep / GMRES_M / buildMatrix / LUSGS with
the reference's serial sums; the device sums dot products in a fixed tree
and assembles the same Jacobian blocks by coloured differences, so the two
agree to the Newton / GMRES tolerances, not bit for bit."""
import os

import numpy as np
import pytest

from ccfd_amd import gpu as G
from ccfd_amd import mesh as M
from oracle import bindings as B
from oracle import timedisc as TD
from tests import cases

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
SOD_IMPLICIT = os.path.join(HERE, "golden", "calc", "sod_implicit.ini")


def _run_pair(mesh, prm, cfg, prim, icfg_g, icfg_o, steps, ptime=1e9,
              tend=1e9):
    gs = G.Solver(mesh, prm)
    orc = B.Oracle(mesh, cfg)
    gs.set_prim(prim)
    orc.set_prim(prim)
    t = 0.0
    ng = no = 0
    for _ in range(steps):
        dt, _ = gs.calc_dt(t, ptime, tend)
        dto, _, _ = orc.calc_dt(t, ptime, tend)
        # states agree to the Newton tolerance, so do the CFL time steps
        assert abs(dt - dto) <= 1e-6 * abs(dto)
        _, nn, gg = gs.implicit_step(icfg_g, t, dt)
        rc, _, nno, ggo = orc.implicit_step(icfg_o, t, dto)
        assert rc == 0
        ng += nn
        no += nno
        t += dt
    out = gs.cons(), orc.cons(), ng, no
    gs.close()
    return out


@pytest.mark.parametrize("precond", [0, 1])
def test_sod_implicit_matches_oracle(precond):
    """calc/riemann/sod_implicit.ini (100 cells, HLLC, CFL 4.99, nKdim 5,
    epsNewton 1e-6): ten backward-Euler steps"""
    mesh, cfg, prim, c = TD.setup_case(SOD_IMPLICIT)
    prm = G.params(gamma=c["gam"], flux=c["flux"], order=c["order"],
                   limiter=c["limiter"], venk_k=c["venk_k"], cfl=c["cfl"],
                   ts1d=int(c["ts1d"]), nrk=c["nRK"], torder=c["timeOrder"])
    ig = G.ImplicitCfg(c["nKdim"], c["nNewtonIter"], precond, c["epsNewton"],
                       c["epsGMRES"], c["gammaEW"])
    io = B.ImplicitCfg(c["nKdim"], c["nNewtonIter"], precond, c["epsNewton"],
                       c["epsGMRES"], c["gammaEW"])
    ug, uo, ng, no = _run_pair(mesh, prm, cfg, prim, ig, io, 10,
                               ptime=c["IOtimeInterval"], tend=c["tEnd"])
    err = cases.rel_err(ug, uo)
    print("precond=%d newton gpu %d oracle %d rel err %.2e" % (precond, ng, no, err))
    assert abs(ng - no) <= 2
    assert err < 1e-6


def test_implicit_second_order_periodic():
    mesh = cases.delaunay_mesh(20, seed=6, periodic=True)
    prim = M.smooth_state(mesh.bary, seed=3)
    prm = G.params(flux=2, order=2, limiter=2, venk_k=5.0, cfl=5.0)
    cfg = B.make_cfg(flux=2, order=2, limiter=2, venk_k=5.0, cfl=5.0)
    ig = G.ImplicitCfg(10, 20, 1, 1e-6, 1e-3, 0.9)
    io = B.ImplicitCfg(10, 20, 1, 1e-6, 1e-3, 0.9)
    ug, uo, ng, no = _run_pair(mesh, prm, cfg, prim, ig, io, 3)
    err = cases.rel_err(ug, uo)
    print("2nd order: newton gpu %d oracle %d rel err %.2e" % (ng, no, err))
    assert err < 1e-7


def test_implicit_fp32_flux_close_to_fp64():
    mesh = cases.delaunay_mesh(20, seed=6, periodic=True)
    prim = M.smooth_state(mesh.bary, seed=3)
    p64 = G.params(flux=5, order=2, limiter=2, venk_k=5.0, cfl=5.0)
    p32 = G.params(flux=5, order=2, limiter=2, venk_k=5.0, cfl=5.0,
                   precision=G.FP32_FLUX)
    ig = G.ImplicitCfg(10, 20, 1, 1e-5, 1e-3, 0.9)
    out = []
    for p in (p64, p32):
        gs = G.Solver(mesh, p)
        gs.set_prim(prim)
        dt, _ = gs.calc_dt(0.0, 1e9, 1e9)
        for k in range(3):
            gs.implicit_step(ig, k * dt, dt)
        out.append(gs.cons())
        gs.close()
    err = cases.rel_err(out[1], out[0])
    print("fp32 vs fp64 implicit rel err %.2e" % err)
    assert err < 1e-4


def test_newton_failure_is_reported():
    """nNewtonIter = 1 cannot converge a large step: the status code says so
    (the reference prints and exits, src/timeDiscretization.c:544-549)"""
    mesh = cases.sod_mesh(100, 1)
    gs = G.Solver(mesh, G.params(flux=5, order=1, cfl=50.0, ts1d=1))
    gs.set_prim(cases.sod_state(mesh))
    dt, _ = gs.calc_dt(0.0, 1.0, 1.0)
    with pytest.raises(G.GpuError, match="Newton NOT converged"):
        gs.implicit_step(G.ImplicitCfg(5, 1, 0, 1e-12, 1e-3, 0.9), 0.0, dt)
    gs.close()


def test_multicolour_lusgs_converges_to_the_same_step():
    """precond = 2 (multicolour sweeps, conservative Jacobian) is another
    preconditioner: the Newton solution of each backward-Euler step is the
    same to the Newton tolerance."""
    mesh, cfg, prim, c = TD.setup_case(SOD_IMPLICIT)
    prm = G.params(gamma=c["gam"], flux=c["flux"], order=c["order"],
                   limiter=c["limiter"], venk_k=c["venk_k"], cfl=c["cfl"],
                   ts1d=int(c["ts1d"]), nrk=c["nRK"], torder=c["timeOrder"])
    out = []
    for pc in (1, 2):
        gs = G.Solver(mesh, prm)
        gs.set_prim(prim)
        t, ng = 0.0, 0
        for _ in range(5):
            dt, _ = gs.calc_dt(t, c["IOtimeInterval"], c["tEnd"])
            _, nn, gg = gs.implicit_step(G.ImplicitCfg(c["nKdim"], 25, pc, 1e-9,
                                                       1e-3, 0.9), t, dt)
            ng += gg
            t += dt
        out.append((gs.cons(), ng))
        gs.close()
    err = cases.rel_err(out[1][0], out[0][0])
    print("reference order vs multicolour: rel err %.2e, GMRES %d / %d"
          % (err, out[0][1], out[1][1]))
    assert err < 1e-7


@pytest.mark.parametrize("precision", [G.FP64, G.FP32_FLUX])
def test_implicit_steady_reduces_the_residual(precision):
    """local pseudo-time steps on an Euler O-grid around a NACA 2412 at
    Mach 0.5: the residual falls by orders of magnitude, and the FP32-flux
    mode (FP64 Newton residual) follows the FP64 solve"""
    raw = M.naca_ogrid(48, 16, wall_bc=101, far_bc=501)
    bcs = [M.BC(101), M.BC(501, rho=1.0, mach=0.5, alpha=2.0, pressure=1.0)]
    mesh = M.Mesh(raw, bcs)
    c = np.sqrt(1.4)
    a = np.radians(2.0)
    prim = np.tile([1.0, 0.5 * c * np.cos(a), 0.5 * c * np.sin(a), 1.0],
                   (mesh.n_elems, 1))
    prm = G.params(flux=2, order=2, limiter=2, venk_k=5.0, precision=precision)
    gs = G.Solver(mesh, prm)
    gs.set_prim(prim)
    r, r0, steps, nn, ng = gs.implicit_steady(G.ImplicitCfg(20, 1, 2, 1e-3,
                                                           1e-2, 0.9),
                                              20.0, 200, 1e-10)
    print("steps %d drop %.2e GMRES %d" % (steps, r[0] / r0[0], ng))
    assert steps == 200 and nn == 200
    assert r[0] / r0[0] < 1e-2
    gs.close()


def _periodic_case():
    mesh = cases.delaunay_mesh(20, seed=6, periodic=True)
    prim = M.smooth_state(mesh.bary, seed=3)
    return mesh, prim


@pytest.mark.parametrize("precond,sweeps,jorder,cycles",
                         [(2, 1, 0, 1), (2, 4, 1, 1), (2, 1, 1, 1),
                          (2, 2, 1, 4), (3, 1, 1, 1), (3, 2, 0, 1)],
                         ids=["lusgs", "sweeps4", "jacobian1", "restarted",
                              "multigrid", "multigrid_jacobian2"])
def test_preconditioner_options_converge_to_the_same_step(precond, sweeps,
                                                          jorder, cycles):
    """Gauss-Seidel sweeps, the multigrid V-cycle, a first-order Jacobian in
    the preconditioner and restarted GMRES change how each Newton system is
    solved, not its solution: three 2nd-order backward-Euler steps agree
    with the oracle's (reference LU-SGS) to the Newton tolerance; on the
    (diagonally dominant) first-order Jacobian more sweeps need fewer
    Krylov iterations"""
    mesh, prim = _periodic_case()
    prm = G.params(flux=2, order=2, limiter=2, venk_k=5.0, cfl=5.0)
    cfg = B.make_cfg(flux=2, order=2, limiter=2, venk_k=5.0, cfl=5.0)
    nk = 4 if cycles > 1 else 10
    ig = G.ImplicitCfg(nk, 40, precond, 1e-10, 1e-3, 0.9, jorder, sweeps,
                       cycles)
    io = B.ImplicitCfg(10, 40, 1, 1e-10, 1e-3, 0.9)
    ug, uo, ng, no = _run_pair(mesh, prm, cfg, prim, ig, io, 3)
    err = cases.rel_err(ug, uo)
    print("precond %d sweeps %d jacobian order %d cycles %d: rel err %.2e, "
          "Newton %d (oracle %d)" % (precond, sweeps, jorder, cycles, err, ng,
                                     no))
    assert err < 1e-6
    if sweeps == 4:
        gs = G.Solver(mesh, prm)
        its = []
        for sw in (1, 4):
            gs.set_prim(prim)
            dt, _ = gs.calc_dt(0.0, 1e9, 1e9)
            _, _, g = gs.implicit_step(G.ImplicitCfg(30, 40, 2, 1e-10, 1e-3,
                                                     0.9, 1, sw, 1), 0.0, dt)
            its.append(g)
        gs.close()
        print("GMRES iterations with 1 / 4 sweeps: %d / %d" % tuple(its))
        assert its[1] < its[0]


def test_implicit_options_are_checked():
    mesh, prim = _periodic_case()
    gs = G.Solver(mesh, G.params(flux=2, order=2, limiter=2, cfl=5.0))
    gs.set_prim(prim)
    dt, _ = gs.calc_dt(0.0, 1e9, 1e9)
    for bad in (dict(jacobianOrder=3), dict(sweeps=65), dict(cycles=101),
                dict(precond=4)):
        ic = G.ImplicitCfg(10, 5, 2, 1e-6, 1e-3, 0.9)
        for k, v in bad.items():
            setattr(ic, k, v)
        with pytest.raises(G.GpuError, match="must be"):
            gs.implicit_step(ic, 0.0, dt)
        with pytest.raises(G.GpuError):
            gs.implicit_steady(ic, 5.0, 3, 1e-6)
    gs.close()


def test_navier_stokes_steady_solve_converges():
    """a coarse NACA 2412 O-grid, Navier-Stokes at Re 500, Mach 0.5: local
    pseudo-time steps with the GMRES-driven CFL control, 4 Gauss-Seidel
    sweeps on the first-order Jacobian; the density residual falls by
    three orders of magnitude in both precisions, and the FP32-flux state
    (FP64 Newton residual) follows the FP64 one to within what is left of
    the transient at that drop"""
    raw = M.naca_ogrid(64, 32, wall_bc=201, far_bc=501, wall_spacing=1e-3,
                       radius=10.0)
    bcs = [M.BC(201), M.BC(501, rho=1.0, mach=0.5, alpha=2.0, pressure=1.0)]
    mesh = M.Mesh(raw, bcs)
    c = np.sqrt(1.4)
    a = np.radians(2.0)
    prim = np.tile([1.0, 0.5 * c * np.cos(a), 0.5 * c * np.sin(a), 1.0],
                   (mesh.n_elems, 1))
    out = []
    for prec in (G.FP64, G.FP32_FLUX):
        prm = G.params(flux=2, order=2, limiter=2, venk_k=5.0, navierstokes=1,
                       mu=0.5 * c / 500.0, precision=prec)
        gs = G.Solver(mesh, prm)
        gs.set_prim(prim)
        ic = G.ImplicitCfg(30, 1, 2, 1e-3, 1e-2, 0.9, 1, 4, 1)
        r, r0, steps, nn, ng = gs.implicit_steady(ic, 1000.0, 400, 1e-3)
        print("%s: steps %d drop %.2e GMRES %d" % (
            "fp64" if prec == G.FP64 else "fp32", steps, r[0] / r0[0], ng))
        assert r[0] <= 1e-3 * r0[0]
        out.append(gs.prim())
        gs.close()
    d = np.linalg.norm(out[1][:, 0] - out[0][:, 0]) / np.linalg.norm(out[0][:, 0])
    print("fp32 vs fp64 density rel L2 %.2e" % d)
    assert d < 1e-2


def _frechet(gs, X, v, time, dt, alpha=1.0):
    """J_F v = v - alpha dt (R(X + eps v) - R(X)) / eps, with R evaluated
    by the device operator at host-set states (independent of the GMRES
    code)"""
    eps = 1e-7 * max(np.linalg.norm(X), 1e-300) / max(np.linalg.norm(v), 1e-300)
    gs.set_cons(X)
    gs.fvtd(time)
    r0 = gs.ut()
    gs.set_cons(X + eps * v)
    gs.fvtd(time)
    r1 = gs.ut()
    gs.set_cons(X)
    return v - alpha * dt * (r1 - r0) / eps


def test_dot_matches_the_oracle():
    mesh, prim = _periodic_case()
    gs = G.Solver(mesh, G.params(flux=2, order=2, limiter=2))
    rng = np.random.default_rng(3)
    a = rng.normal(size=(mesh.n_elems, 4))
    b = rng.normal(size=(mesh.n_elems, 4))
    d, o = gs.dot(a, b), B.dot(a, b)
    print("device %.17g oracle %.17g" % (d, o))
    assert abs(d - o) <= 1e-13 * np.abs(a * b).sum()
    assert gs.dot(a, a) == gs.dot(a, a)     # fixed summation order
    gs.close()


@pytest.mark.parametrize("precond", [0, 1, 2, 3])
def test_gmres_solves_the_linear_system(precond):
    """GMRES_M through the C ABI: |B + J_F dX| <= epsGMRES |B| with J_F
    applied by an independent finite difference, and dX -> -B as dt -> 0"""
    mesh, prim = _periodic_case()
    prm = G.params(flux=2, order=2, limiter=2, venk_k=5.0, cfl=5.0)
    gs = G.Solver(mesh, prm)
    gs.set_prim(prim)
    X = gs.cons()
    dt, _ = gs.calc_dt(0.0, 1e9, 1e9)
    gs.set_cons(X)
    rng = np.random.default_rng(11)
    Bv = rng.normal(size=X.shape) * 1e-3
    cfg = G.ImplicitCfg(150, 1, precond, 1e-6, 1e-6, 0.9)
    dX, it = gs.gmres(cfg, 0.0, 0.0, dt, 1.0, Bv)
    res = np.linalg.norm(Bv + _frechet(gs, X, dX, 0.0, dt))
    print("precond %d: %d iterations, |B + J dX| / |B| = %.2e"
          % (precond, it, res / np.linalg.norm(Bv)))
    assert 0 < it < 150       # converged before the Krylov space ran out
    assert res <= 1e-5 * np.linalg.norm(Bv)
    dX0, _ = gs.gmres(cfg, 0.0, 0.0, 1e-12 * dt, 1.0, Bv)
    assert np.allclose(dX0, -Bv, rtol=1e-8, atol=1e-12)
    gs.close()


def test_multigrid_needs_fewer_krylov_iterations():
    """a hard system, (I - dt J) with dt a thousand times the CFL step on a
    128x64 NACA Navier-Stokes O-grid (three coarse levels): one V-cycle per
    preconditioner application leaves GMRES fewer iterations to reach 1e-6
    than the Gauss-Seidel sweeps it uses as smoother"""
    raw = M.naca_ogrid(128, 64, wall_bc=201, far_bc=501, wall_spacing=1e-3,
                       radius=10.0)
    bcs = [M.BC(201), M.BC(501, rho=1.0, mach=0.5, alpha=2.0, pressure=1.0)]
    mesh = M.Mesh(raw, bcs)
    c = np.sqrt(1.4)
    a = np.radians(2.0)
    prim = np.tile([1.0, 0.5 * c * np.cos(a), 0.5 * c * np.sin(a), 1.0],
                   (mesh.n_elems, 1))
    prm = G.params(flux=2, order=1, navierstokes=1, mu=0.5 * c / 500.0,
                   cfl=1.0)
    rng = np.random.default_rng(5)
    Bv = rng.normal(size=(mesh.n_elems, 4))
    its = {}
    for pc in (2, 3):
        gs = G.Solver(mesh, prm)
        gs.set_prim(prim)
        dt, _ = gs.calc_dt(0.0, 1e9, 1e9)
        _, its[pc] = gs.gmres(G.ImplicitCfg(200, 1, pc, 1e-6, 1e-6, 0.9),
                              0.0, 0.0, 1e3 * dt, 1.0, Bv)
        gs.close()
    print("GMRES iterations to 1e-6: Gauss-Seidel %d, multigrid %d"
          % (its[2], its[3]))
    assert its[3] < its[2]


def test_steady_solve_continues_across_calls():
    """a steady solve split into calls (res0 passed back) is the same solve:
    20 + 20 steps give the state and residuals of 40 steps in one call"""
    raw = M.naca_ogrid(48, 16, wall_bc=101, far_bc=501)
    bcs = [M.BC(101), M.BC(501, rho=1.0, mach=0.5, alpha=2.0, pressure=1.0)]
    mesh = M.Mesh(raw, bcs)
    c = np.sqrt(1.4)
    a = np.radians(2.0)
    prim = np.tile([1.0, 0.5 * c * np.cos(a), 0.5 * c * np.sin(a), 1.0],
                   (mesh.n_elems, 1))
    prm = G.params(flux=2, order=2, limiter=2, venk_k=5.0)
    ic = G.ImplicitCfg(20, 1, 2, 1e-3, 1e-2, 0.9, 1, 2, 1)
    out = []
    for split in (False, True):
        gs = G.Solver(mesh, prm)
        gs.set_prim(prim)
        if split:
            r, r0, n1, _, _ = gs.implicit_steady(ic, 100.0, 20, 1e-12)
            r, r0b, n2, _, _ = gs.implicit_steady(ic, 100.0, 20, 1e-12, r0)
            assert n1 == n2 == 20 and np.array_equal(r0, r0b)
        else:
            r, r0, n, _, _ = gs.implicit_steady(ic, 100.0, 40, 1e-12)
            assert n == 40
        out.append((gs.cons(), r, r0))
        gs.close()
    print("drop after 40 steps: %.3e" % (out[0][1][0] / out[0][2][0]))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    assert np.array_equal(out[0][2], out[1][2])


def _local_step_reference(mesh, cfg, prim0, cfl, gam=1.4):
    """the backward-Euler pseudo-step with local steps, X - Q - dt_i R(X) = 0,
    solved by Newton with a dense finite-difference Jacobian; R from the
    oracle, dt_i the CFL limit of calcTimeStep per element
    (src/timeDiscretization.c:253-271) at the initial state"""
    from ccfd_amd import serve as SV
    orc = B.Oracle(mesh, cfg)
    p = prim0
    a = np.sqrt(gam * p[:, 3] / p[:, 0])
    dt = cfl * mesh.area / ((np.abs(p[:, 1]) + a) * mesh.sx
                            + (np.abs(p[:, 2]) + a) * mesh.sy)
    orc.set_prim(prim0)
    Q = orc.cons().copy()

    def R(X):
        orc.set_prim(SV.cons_to_prim(X.reshape(-1, 4), gam))
        return orc.fvtd(0.0).reshape(-1)

    def F(X):
        return X - Q.reshape(-1) - np.repeat(dt, 4) * R(X)

    X = Q.reshape(-1).copy()
    f0 = np.linalg.norm(F(X))
    for _ in range(8):
        fx = F(X)
        if np.linalg.norm(fx) <= 1e-13 * f0:
            break
        J = np.empty((X.size, X.size))
        for j in range(X.size):
            h = 1e-7 * max(1.0, abs(X[j]))
            Xp = X.copy()
            Xp[j] += h
            J[:, j] = (F(Xp) - fx) / h
        X = X - np.linalg.solve(J, fx)
    assert np.linalg.norm(F(X)) <= 1e-12 * f0
    return X.reshape(-1, 4), Q


def test_steady_pseudo_step_solves_local_backward_euler():
    """one pseudo-step of the steady driver, its Newton loop converged
    tightly, lands on the solution of X - Q - dt_i R(X) = 0 with the local
    steps dt_i of calcTimeStep's CFL limit (CFL 1 on the first step), as an
    independent dense Newton solve on the oracle's operator finds it: the
    local steps, the row scaling (which must not move the solution) and
    the Newton loop of the steady path, pinned at once"""
    raw = M.naca_ogrid(32, 12, wall_bc=101, far_bc=501)
    bcs = [M.BC(101), M.BC(501, rho=1.0, mach=0.5, alpha=2.0, pressure=1.0)]
    mesh = M.Mesh(raw, bcs)
    c = np.sqrt(1.4)
    al = np.radians(2.0)
    prim = np.tile([1.0, 0.5 * c * np.cos(al), 0.5 * c * np.sin(al), 1.0],
                   (mesh.n_elems, 1))
    prim += 0.02 * M.smooth_state(mesh.bary, seed=9) - 0.02 * prim
    kw = dict(flux=2, order=2, limiter=2, venk_k=1000.0)
    X, Q = _local_step_reference(mesh, B.make_cfg(**kw), prim, 1.0)
    gs = G.Solver(mesh, G.params(**kw))
    gs.set_prim(prim)
    assert np.allclose(gs.cons(), Q, rtol=1e-14, atol=1e-15)
    ic = G.ImplicitCfg(60, 30, 2, 1e-11, 1e-8, 0.9, 1, 2, 3)
    _, _, steps, nn, _ = gs.implicit_steady(ic, 1.0, 1, 1e-30)
    err = cases.rel_err(gs.cons(), X)
    print("Newton iterations %d, rel err vs the dense solve %.2e, step size "
          "%.2e" % (nn, err, cases.rel_err(X, Q)))
    assert steps == 1
    assert err < 1e-7
    assert cases.rel_err(X, Q) > 1e-3      # the step moved the state
    gs.close()


def _naca_ns_case(ni=32, nj=12):
    raw = M.naca_ogrid(ni, nj, wall_bc=201, far_bc=501, wall_spacing=1e-3,
                       radius=20.0)
    bcs = [M.BC(201), M.BC(501, rho=1.0, mach=0.5, alpha=2.0, pressure=1.0)]
    mesh = M.Mesh(raw, bcs)
    c = np.sqrt(1.4)
    a = np.radians(2.0)
    prim = np.tile([1.0, 0.5 * c * np.cos(a), 0.5 * c * np.sin(a), 1.0],
                   (mesh.n_elems, 1))
    prim += 0.02 * (M.smooth_state(mesh.bary, seed=9) - prim)
    return mesh, prim


@pytest.mark.parametrize("case", ["naca_ns", "delaunay_euler", "mixed_bc"])
@pytest.mark.parametrize("precision", [G.FP64, G.FP32_FLUX])
def test_elementwise_jacobian_equals_the_coloured_differences(case,
                                                             precision):
    """the first-order conservative Jacobian of the preconditioner
    (precond 2, jacobianOrder 1 on a 2nd-order operator), assembled element
    by element in one launch, equals the coloured finite differences bit for
    bit: same perturbations, same side fluxes, same sums (Navier-Stokes
    walls and far field, periodic triangles, every boundary type)"""
    if case == "naca_ns":
        mesh, prim = _naca_ns_case()
        kw = dict(flux=2, order=2, limiter=2, venk_k=5.0, navierstokes=1,
                  mu=0.5 * np.sqrt(1.4) / 5000.0)
    elif case == "delaunay_euler":
        mesh, prim = _periodic_case()
        kw = dict(flux=5, order=2, limiter=2, venk_k=5.0)
    else:
        mesh, _ = cases.mixed_bc_mesh()
        prim = cases.random_state(mesh.n_elems, seed=4, mach=0.6)
        kw = dict(flux=9, order=2, limiter=1)
    gs = G.Solver(mesh, G.params(precision=precision, **kw))
    gs.set_prim(prim)
    dt, _ = gs.calc_dt(0.0, 1e9, 1e9)
    cfg = G.ImplicitCfg(10, 1, 2, 1e-3, 1e-2, 0.9, 1, 1, 1)
    D0, B0 = gs.jacobian_blocks(cfg, 0.0, 10 * dt, method=0)
    D1, B1 = gs.jacobian_blocks(cfg, 0.0, 10 * dt, method=1)
    gs.close()
    off = np.abs(B0).max()
    print("%s: |D - I| max %.3e, side blocks max %.3e"
          % (case, np.abs(D0 - np.eye(4)).max(), off))
    assert off > 0.0 and np.abs(D0 - np.eye(4)).max() > 0.0
    assert np.array_equal(D0, D1)
    assert np.array_equal(B0, B1)


def test_build_matrix_matches_the_oracle():
    """buildMatrix + calcDinv with the reference's primitive perturbation
    (precond 1, src/linearSolver.c:186-252) at 2nd order on the every-BC
    mesh: the device's coloured differences record the blocks the oracle's
    element-by-element differences record (the two perturb different
    sets of elements at once, which the distance-3 colouring makes
    equivalent; the base states differ by rounding only, which the
    difference quotient amplifies to ~1e-8)"""
    mesh, _ = cases.mixed_bc_mesh()
    prim = cases.random_state(mesh.n_elems, seed=4, mach=0.6)
    kw = dict(flux=5, order=2, limiter=2, venk_k=5.0)
    gs = G.Solver(mesh, G.params(**kw))
    gs.set_prim(prim)
    # the state as a Newton iteration holds it: consPrim of the
    # conservative state
    gs.set_cons(gs.cons())
    prim2 = gs.prim()
    dt = 0.01
    D, Bk = gs.jacobian_blocks(G.ImplicitCfg(10, 1, 1, 1e-3, 1e-2, 0.9), 0.3,
                               dt, method=0)
    gs.close()
    orc = B.Oracle(mesh, B.make_cfg(**kw))
    orc.set_prim(prim2)
    Do, Bo, _ = orc.build_matrix(0.3, dt)
    sd = np.abs(Do - np.eye(4)).max()
    sb = np.abs(Bo).max()
    ed = np.abs(D - Do).max() / sd
    eb = np.abs(Bk - Bo).max() / sb
    print("D rel err %.2e, side blocks rel err %.2e" % (ed, eb))
    assert ed < 1e-6 and eb < 1e-6
