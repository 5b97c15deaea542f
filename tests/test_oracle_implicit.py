"""Pins the oracle's implicit step (implicitTimeStep + GMRES_M + buildMatrix +
LUSGS, src/timeDiscretization.c:412-552, src/linearSolver.c:136-510) by what
the mathematics fixes, since the reference's timeDiscretization.c and
linearSolver.c cannot be compiled here (they include the CGNS headers):

* the converged step solves backward Euler: F(U1) = U1 - U0 - dt R(U1) is
  small, with R evaluated by a separate operator call on U1;
* as dt -> 0 the step approaches explicit Euler, U0 + dt R(U0), to O(dt^2);
* LU-SGS preconditioning changes the Krylov path, not the converged answer,
  and cuts the GMRES iterations;
* a uniform state on a periodic mesh is a fixed point (R = 0).

A wrong sign or factor in the Newton function, the Frechet difference, the
Jacobian assembly or the preconditioner sweeps fails one of these.
"""
import numpy as np
import pytest

from ccfd_amd import mesh as M
from oracle import bindings as B
from tests import cases

GAM = 1.4


def _cons(prim):
    c = np.empty_like(prim)
    c[:, 0] = prim[:, 0]
    c[:, 1] = prim[:, 0] * prim[:, 1]
    c[:, 2] = prim[:, 0] * prim[:, 2]
    c[:, 3] = prim[:, 3] / (GAM - 1.0) + 0.5 * prim[:, 0] * (
        prim[:, 1] ** 2 + prim[:, 2] ** 2)
    return c


def _rhs(mesh, cfg, prim):
    o = B.Oracle(mesh, cfg)
    o.set_prim(prim)
    return o.fvtd(0.0)


def _case(order):
    mesh = cases.delaunay_mesh(16, seed=5, periodic=True)
    prim = M.smooth_state(mesh.bary, seed=4)
    cfg = B.make_cfg(flux=5, order=order, limiter=2, venk_k=5.0, cfl=4.0)
    return mesh, prim, cfg


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("precond", [0, 1])
def test_step_solves_backward_euler(order, precond):
    mesh, prim, cfg = _case(order)
    o = B.Oracle(mesh, cfg)
    o.set_prim(prim)
    dt, _, _ = o.calc_dt(0.0, 1e9, 1e9)
    u0 = o.cons()
    ic = B.ImplicitCfg(30, 30, precond, 1e-10, 1e-6, 0.9)
    rc, _, nn, ng = o.implicit_step(ic, 0.0, dt)
    assert rc == 0 and nn >= 1 and ng >= 1
    u1 = o.cons()
    f0 = -dt * _rhs(mesh, cfg, prim)                 # F(U0)
    f1 = u1 - u0 - dt * _rhs(mesh, cfg, o.prim())    # F(U1)
    ratio = np.linalg.norm(f1) / np.linalg.norm(f0)
    print("order %d precond %d: newton %d gmres %d |F1|/|F0| %.2e"
          % (order, precond, nn, ng, ratio))
    assert ratio < 1e-7
    # and the step moved the state by about dt R (CFL 4: not a tiny step)
    assert np.linalg.norm(u1 - u0) > 0.1 * np.linalg.norm(f0)


def test_small_dt_limit_is_explicit_euler():
    mesh, prim, cfg = _case(1)
    o = B.Oracle(mesh, cfg)
    o.set_prim(prim)
    dt, _, _ = o.calc_dt(0.0, 1e9, 1e9)
    errs = []
    for scale in (1e-2, 1e-3):
        h = dt * scale
        o.set_prim(prim)
        rc, _, _, _ = o.implicit_step(B.ImplicitCfg(30, 30, 1, 1e-9, 1e-8, 0.9),
                                      0.0, h)
        assert rc == 0
        ue = _cons(prim) + h * _rhs(mesh, cfg, prim)
        du = np.linalg.norm(ue - _cons(prim))
        errs.append(np.linalg.norm(o.cons() - ue) / du)
    print("relative distance to explicit Euler:", errs)
    # first-order agreement: the gap shrinks with dt (O(dt^2) / O(dt))
    assert errs[0] < 5e-2
    assert errs[1] < 0.2 * errs[0]


def test_preconditioner_keeps_the_answer():
    mesh, prim, cfg = _case(2)
    out = []
    for precond in (0, 1):
        o = B.Oracle(mesh, cfg)
        o.set_prim(prim)
        dt, _, _ = o.calc_dt(0.0, 1e9, 1e9)
        rc, _, nn, ng = o.implicit_step(
            B.ImplicitCfg(30, 30, precond, 1e-10, 1e-6, 0.9), 0.0, dt)
        assert rc == 0
        out.append((o.cons(), ng))
    scale = np.abs(out[0][0]).max(axis=0)
    assert (np.abs(out[0][0] - out[1][0]) / scale).max() < 1e-8
    print("GMRES iterations without / with LU-SGS:", out[0][1], out[1][1])
    assert out[1][1] < out[0][1]


def test_uniform_state_is_fixed_point():
    mesh = cases.delaunay_mesh(12, seed=2, periodic=True)
    prim = np.tile([1.2, 0.3, -0.2, 0.9], (mesh.n_elems, 1))
    cfg = B.make_cfg(flux=2, order=2, limiter=2, venk_k=5.0, cfl=4.0)
    o = B.Oracle(mesh, cfg)
    o.set_prim(prim)
    u0 = o.cons()
    dt, _, _ = o.calc_dt(0.0, 1e9, 1e9)
    rc, _, _, _ = o.implicit_step(B.ImplicitCfg(10, 10, 1, 1e-6, 1e-3, 0.9),
                                  0.0, dt)
    assert rc == 0
    np.testing.assert_allclose(o.cons(), u0, rtol=1e-14, atol=1e-14)
