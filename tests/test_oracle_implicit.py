"""Pins the oracle's implicit step (implicitTimeStep + GMRES_M + buildMatrix +
LUSGS, src/timeDiscretization.c:412-552, src/linearSolver.c:136-510) by what
the mathematics fixes, since the reference's timeDiscretization.c and
linearSolver.c cannot be compiled here (they include the CGNS headers):

* the converged step solves backward Euler: F(U1) = U1 - U0 - dt R(U1) is
  small, with R evaluated by a separate operator call on U1;
* as dt -> 0 the step approaches explicit Euler, U0 + dt R(U0), to O(dt^2);
* LU-SGS preconditioning changes the Krylov path, not the converged answer,
  and cuts the GMRES iterations;
* a uniform state on a periodic mesh is a fixed point (R = 0);
* the Jacobian blocks (buildMatrix) are I - dt dR/dpVar: the column of one
  primitive variable of one element is what a forward difference of R,
  from separate operator calls with a step of its own, gives in that
  element and its face neighbours, and the stored inverses invert the
  diagonal blocks;
* LU-SGS returns x with (D + L) D^-1 (D + U) x = b for the blocks above,
  the product assembled here with scipy from the blocks by element index;
* GMRES_M with a full Krylov space returns the dense solution of
  (I - dt dR/dU) x = -b, the matrix built here from forward differences of
  R, with and without LU-SGS.

A wrong sign or factor in the Newton function, the Frechet difference, the
Jacobian assembly or the preconditioner sweeps fails one of these; a
transposed block or swapped lower and upper halves fail the last two, which
the converged steps alone would not show (the preconditioner changes the
Krylov path, not the answer).
"""
import numpy as np
import pytest
import scipy.sparse as sp

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


def _block_case(kind, order):
    if kind == "periodic":
        mesh = cases.delaunay_mesh(12, seed=5, periodic=True)
    else:
        mesh, _ = cases.mixed_bc_mesh(12, 6)
    prim = M.smooth_state(mesh.bary, seed=3)
    cfg = B.make_cfg(flux=5 if order == 2 else 2, order=order, limiter=2,
                     venk_k=5.0, cfl=4.0)
    own, nbr = mesh.slot_elem(), mesh.nbr
    inner = nbr >= 0
    # one slot per (element, neighbour) pair and no self-neighbours, so
    # every stored block is one entry of the dense matrix
    assert not np.any(nbr[inner] == own[inner])
    pairs = own[inner] * mesh.n_elems + nbr[inner]
    assert len(np.unique(pairs)) == len(pairs)
    return mesh, prim, cfg, own, nbr


def _bsr(blocks, rows, cols, n):
    """sparse 4n x 4n matrix from 4x4 blocks at (rows, cols)"""
    a = np.arange(4)
    r = np.broadcast_to(4 * rows[:, None, None] + a[None, :, None], blocks.shape)
    c = np.broadcast_to(4 * cols[:, None, None] + a[None, None, :], blocks.shape)
    return sp.csr_matrix((blocks.ravel(), (r.ravel(), c.ravel())),
                         shape=(4 * n, 4 * n))


@pytest.mark.parametrize("kind", ["periodic", "walls"])
@pytest.mark.parametrize("order", [1, 2])
def test_jacobian_blocks_are_the_derivative_of_R(kind, order):
    mesh, prim, cfg, own, nbr = _block_case(kind, order)
    o = B.Oracle(mesh, cfg)
    o.set_prim(prim)
    dt, _, _ = o.calc_dt(0.0, 1e9, 1e9)
    Dg, Blk, Dinv = o.build_matrix(0.0, dt)
    np.testing.assert_allclose(Dinv, np.linalg.inv(Dg), rtol=1e-10,
                               atol=1e-12 * np.abs(Dinv).max())
    # forward differences, one primitive component at a time, with a step
    # of their own: R is not differentiable at boundary states whose ghost
    # copies or mirrors the interior (a central difference there misses by
    # 6e-2), and the reference's finite differences record the forward slope
    rng = np.random.default_rng(order)
    h = 1e-7
    o.set_prim(prim)
    r0 = o.fvtd(0.0)
    worst = 0.0
    for k in rng.choice(mesh.n_elems, 8, replace=False):
        nb = np.nonzero(nbr == k)[0]            # slots of rows j, column k
        assert len(nb) >= 2
        for iv in range(4):
            p = prim.copy()
            p[k, iv] += h
            o.set_prim(p)
            jd = (o.fvtd(0.0) - r0) / h         # dR/dpVar[k, iv], every row
            e = np.eye(4)[iv]
            want = {k: e - dt * jd[k]}
            got = {k: Dg[k, :, iv]}
            for s in nb:
                want[own[s]] = -dt * jd[own[s]]
                got[own[s]] = Blk[s, :, iv]
            if order == 1:
                # first order: R reaches the face neighbours only
                far = np.setdiff1d(np.arange(mesh.n_elems), list(want))
                assert np.abs(jd[far]).max() == 0.0
            scale = max(np.abs(w).max() for w in want.values())
            err = max(np.abs(got[j] - want[j]).max() for j in want) / scale
            worst = max(worst, err)
    print("%s order %d: worst block error %.2e" % (kind, order, worst))
    assert worst < 1e-5


@pytest.mark.parametrize("kind", ["periodic", "walls"])
def test_lusgs_inverts_the_symmetric_gauss_seidel_product(kind):
    mesh, prim, cfg, own, nbr = _block_case(kind, 2)
    o = B.Oracle(mesh, cfg)
    o.set_prim(prim)
    dt, _, _ = o.calc_dt(0.0, 1e9, 1e9)
    Dg, Blk, Dinv = o.build_matrix(0.0, dt)
    n = mesh.n_elems
    b = np.random.default_rng(9).normal(size=(n, 4))
    x = o.lusgs(Dinv, Blk, b)
    diag = np.arange(n)
    lo = (nbr >= 0) & (nbr < own)
    up = nbr > own
    D = _bsr(Dg, diag, diag, n)
    Di = _bsr(np.linalg.inv(Dg), diag, diag, n)
    L = _bsr(Blk[lo], own[lo], nbr[lo], n)
    U = _bsr(Blk[up], own[up], nbr[up], n)
    r = (D + L) @ (Di @ ((D + U) @ x.ravel())) - b.ravel()
    rel = np.linalg.norm(r) / np.linalg.norm(b)
    print("%s: |(D+L) D^-1 (D+U) x - b| / |b| = %.2e" % (kind, rel))
    assert rel < 1e-12
    # and the sweeps are not block Jacobi: the off-diagonal blocks matter
    assert np.linalg.norm(D @ x.ravel() - b.ravel()) > 1e-3 * np.linalg.norm(b)


def _prim_of(cons):
    p = np.empty_like(cons)
    p[:, 0] = cons[:, 0]
    p[:, 1] = cons[:, 1] / cons[:, 0]
    p[:, 2] = cons[:, 2] / cons[:, 0]
    p[:, 3] = (GAM - 1.0) * (cons[:, 3] - 0.5 * (cons[:, 1] ** 2 +
                                                 cons[:, 2] ** 2) / cons[:, 0])
    return p


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("precond", [0, 1])
def test_gmres_solves_the_linear_system(order, precond):
    """GMRES_M with a Krylov space as large as the system and a tight
    tolerance returns the solution of (I - dt dR/dU) x = -b, the matrix
    assembled here column by column from forward differences of R in the
    conservative variables (separate operator calls, a step of its own):
    a wrong sign, Givens rotation, back-substitution or right
    preconditioning (x from Z, not V) fails it"""
    mesh = cases.delaunay_mesh(8, seed=3, periodic=True)
    prim = M.smooth_state(mesh.bary, seed=6)
    cfg = B.make_cfg(flux=5, order=order, limiter=2, venk_k=5.0, cfl=2.0)
    o = B.Oracle(mesh, cfg)
    o.set_prim(prim)
    dt, _, _ = o.calc_dt(0.0, 1e9, 1e9)
    u0 = o.cons()
    n = 4 * mesh.n_elems
    b = np.random.default_rng(order + 2 * precond).normal(size=(mesh.n_elems, 4))
    ic = B.ImplicitCfg(n, 1, precond, 1e-3, 1e-11, 0.9)
    x, iters = o.gmres(ic, 0.0, dt, b)
    assert np.array_equal(o.prim(), prim)            # state left as it was
    # dense I - dt dR/dU by forward differences at the operator time dt
    r0 = _rhs_at(mesh, cfg, prim, dt)
    A = np.eye(n)
    h = 1e-7
    for j in range(n):
        u = u0.copy()
        u.flat[j] += h * max(1.0, abs(u0.flat[j]))
        step = u.flat[j] - u0.flat[j]
        A[:, j] -= dt * ((_rhs_at(mesh, cfg, _prim_of(u), dt) - r0) / step).ravel()
    want = np.linalg.solve(A, -b.ravel())
    err = np.linalg.norm(x.ravel() - want) / np.linalg.norm(want)
    res = np.linalg.norm(A @ x.ravel() + b.ravel()) / np.linalg.norm(b)
    print("order %d precond %d: %d iterations, |x - x_dense| / |x| %.1e, "
          "residual %.1e" % (order, precond, iters, err, res))
    assert err < 1e-5 and res < 1e-5
    assert iters < n


def _rhs_at(mesh, cfg, prim, time):
    o = B.Oracle(mesh, cfg)
    o.set_prim(prim)
    return o.fvtd(time)


def test_dot_is_the_sum_of_products():
    """vectorDotProduct (src/linearSolver.c:115-128): within a few ulps of
    the correctly rounded sum (math.fsum) of the elementwise products, on
    vectors whose terms cancel"""
    import math
    rng = np.random.default_rng(12)
    for n in (1, 7, 1000, 40000):
        a = rng.normal(size=(n, 4)) * np.exp(rng.uniform(-5, 5, (n, 1)))
        b = rng.normal(size=(n, 4))
        exact = math.fsum((a * b).ravel().tolist())
        scale = math.fsum(np.abs(a * b).ravel().tolist())
        assert abs(B.dot(a, b) - exact) <= 1e-15 * n * scale
    assert B.dot(np.zeros((3, 4)), np.ones((3, 4))) == 0.0


@pytest.mark.parametrize("precond,newton,gmres", [(1, 60, 300),
                                                  (0, 107, 535)])
def test_sod_implicit_run_takes_the_reference_iterations(precond, newton,
                                                          gmres):
    """calc/riemann/sod_implicit.ini run to tEnd: 10 time steps with the
    Newton and GMRES totals the reference's own run of that file prints
    (src/timeDiscretization.c:755-756): 60 / 300 with LU-SGS, 107 / 535
    without. Each Newton count depends on every residual norm of the run,
    so a change in the step, the Jacobian, the preconditioner or the
    Eisenstat-Walker tolerance shows here."""
    import os
    from oracle import timedisc as TD
    ini = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                       "calc", "sod_implicit.ini")
    mesh, cfg, prim, c = TD.setup_case(ini)
    orc = B.Oracle(mesh, cfg)
    orc.set_prim(prim)
    ic = B.ImplicitCfg(c["nKdim"], c["nNewtonIter"], precond, c["epsNewton"],
                       c["epsGMRES"], c["gammaEW"])
    t, steps, nn, ng = 0.0, 0, 0, 0
    while c["tEnd"] - t > 1e-15:
        dt, _, bad = orc.calc_dt(t, c["IOtimeInterval"], c["tEnd"])
        assert not bad
        rc, _, a, b = orc.implicit_step(ic, t, dt)
        assert rc == 0
        nn, ng, t, steps = nn + a, ng + b, t + dt, steps + 1
    assert (steps, nn, ng) == (10, newton, gmres)
