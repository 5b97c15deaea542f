"""Host build of the device physics header (ccfd_amd/csrc/physics.hpp,
compiled by build() into tests/native/libphysics_host.so) against the
reference's flux vectors, bit for bit, and the boundary ghost states against
what each boundary type fixes. Checks the transcription of the HIP physics
code without a GPU."""
import ctypes as C
import os

import numpy as np
import pytest

from tests import golden_io

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native",
                   "libphysics_host.so")
dp = C.POINTER(C.c_double)


def _lib():
    if not os.path.exists(LIB):
        pytest.skip("physics host library not built")
    return C.CDLL(LIB)


@pytest.mark.parametrize("code", list(range(1, 12)))
def test_device_physics_matches_reference_fluxes(code):
    L = _lib()
    fx = golden_io.flux_vectors()
    l, r = fx["L"], fx["R"]
    F = np.empty_like(l)
    L.phys_riemann(code, 0, C.c_double(1.4), len(l), l.ctypes.data_as(dp),
                   r.ctypes.data_as(dp), F.ctypes.data_as(dp))
    assert np.array_equal(F, fx["F%02d" % code], equal_nan=True)


@pytest.mark.parametrize("code", list(range(1, 12)))
def test_fp32_fluxes_close_to_fp64(code):
    L = _lib()
    fx = golden_io.flux_vectors()
    l, r = fx["L"][:1000], fx["R"][:1000]
    F = np.empty_like(l)
    L.phys_riemann(code, 1, C.c_double(1.4), len(l), l.ctypes.data_as(dp),
                   r.ctypes.data_as(dp), F.ctypes.data_as(dp))
    ref = fx["F%02d" % code][:1000]
    scale = np.abs(ref).max(axis=1, keepdims=True) + 1e-30
    err = (np.abs(F - ref) / scale).max(axis=1)
    # states span Mach 0..3 and density / pressure ratios up to e^6: single
    # precision loses a few digits only in cancellation-heavy corners
    assert np.median(err) < 1e-6
    assert err.max() < 5e-4


def _bc(L, btype, bcp, interior, normals, mu=0.0):
    out = np.empty_like(interior)
    bcp = np.ascontiguousarray(bcp, dtype=np.float64)
    L.phys_bc(btype, bcp.ctypes.data_as(dp), C.c_double(1.4), C.c_double(mu),
              len(interior), normals.ctypes.data_as(dp),
              interior.ctypes.data_as(dp), out.ctypes.data_as(dp))
    return out


def test_boundary_states_closed_forms():
    """the device ghost states (bc_ghost, src/boundary.c:147-294) against
    what each boundary type fixes: slip walls mirror the normal velocity,
    no-slip walls reverse it, inflow takes the boundary state, outflow the
    interior, the pressure outlet sets the boundary pressure below sonic
    outflow (density scaled by the pressure ratio), and the characteristic
    boundary reduces to the interior (supersonic outflow), to the boundary
    state (supersonic inflow) and to a fixed point where both agree"""
    L = _lib()
    rng = np.random.default_rng(4)
    n = 200
    q = np.column_stack([rng.uniform(0.2, 2, n), rng.uniform(-1, 1, n),
                         rng.uniform(-1, 1, n), rng.uniform(0.2, 2, n)])
    th = rng.uniform(0, 2 * np.pi, n)
    nrm = np.ascontiguousarray(np.column_stack([np.cos(th), np.sin(th)]))
    vn = (q[:, 1:3] * nrm).sum(axis=1)
    mirror = q[:, 1:3] - 2 * vn[:, None] * nrm

    g = _bc(L, 1, np.zeros(4), q, nrm)
    assert np.array_equal(g[:, [0, 3]], q[:, [0, 3]])
    np.testing.assert_allclose(g[:, 1:3], mirror, rtol=0, atol=1e-15)
    g = _bc(L, 2, np.zeros(4), q, nrm, mu=1e-3)
    np.testing.assert_allclose(g[:, 1:3], -q[:, 1:3], rtol=0, atol=1e-15)
    g = _bc(L, 2, np.zeros(4), q, nrm, mu=0.0)     # inviscid: a slip wall
    np.testing.assert_allclose(g[:, 1:3], mirror, rtol=0, atol=1e-15)

    state = np.array([1.1, 0.3, -0.2, 0.9])
    assert np.array_equal(_bc(L, 3, state, q, nrm), np.tile(state, (n, 1)))
    assert np.array_equal(_bc(L, 4, state, q, nrm), q)

    g = _bc(L, 8, np.array([0.0, 0.0, 0.0, 0.7]), q, nrm)
    c = np.sqrt(1.4 * q[:, 3] / q[:, 0])
    sub = vn / c < 1.0
    assert sub.any() and (~sub).any()       # both branches occur
    np.testing.assert_allclose(g[~sub], q[~sub], rtol=1e-15, atol=0)
    np.testing.assert_allclose(g[sub, 3], 0.7, rtol=0, atol=0)
    np.testing.assert_allclose(g[sub, 0], q[sub, 0] * 0.7 / q[sub, 3],
                               rtol=1e-15)
    assert np.array_equal(g[:, 1:3], q[:, 1:3])

    # characteristic boundary: limits of the eigenvalue test
    for k in range(20):
        nk = nrm[k:k + 1]
        cb = np.sqrt(1.4 * state[3] / state[0])
        for speed, want in ((2.5 * cb, "interior"), (-2.5 * cb, "boundary")):
            bc = state.copy()
            bc[1:3] = speed * nk[0] + 0.1 * np.array([-nk[0, 1], nk[0, 0]])
            g = _bc(L, 5, bc, q[k:k + 1].copy(), nk)
            ref = q[k] if want == "interior" else bc
            np.testing.assert_allclose(g[0], ref, rtol=1e-13, atol=1e-13)
        g = _bc(L, 5, q[k], q[k:k + 1].copy(), nk)
        np.testing.assert_allclose(g[0], q[k], rtol=1e-13, atol=1e-13)


def _conv(L, mode, x, ref=None, width=4):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros((len(x), width))
    r = np.ascontiguousarray(ref if ref is not None else np.zeros((len(x), 4)))
    L.phys_conv(mode, C.c_double(1.4), len(x), x.ctypes.data_as(dp),
                out.ctypes.data_as(dp), r.ctypes.data_as(dp))
    return out


def test_state_conversions():
    """equation of state and characteristic transform of the device code
    (src/equationOfState.c:19-110): primitive <-> conservative round trip;
    the positivity clamp of consPrim; char_cons inverts cons_char; and the
    columns of char_cons are right eigenvectors of the 1-D Euler flux
    Jacobian at the reference state with eigenvalues u - c, u, u + c"""
    L = _lib()
    rng = np.random.default_rng(8)
    n = 100
    p = np.column_stack([rng.uniform(0.1, 3, n), rng.uniform(-2, 2, n),
                         rng.uniform(-2, 2, n), rng.uniform(0.1, 3, n)])
    c = _conv(L, 0, p)
    assert np.allclose(c[:, 3], p[:, 3] / 0.4 + 0.5 * p[:, 0] *
                       (p[:, 1] ** 2 + p[:, 2] ** 2), rtol=1e-14)
    np.testing.assert_allclose(_conv(L, 1, c), p, rtol=1e-13, atol=1e-14)
    # clamp: vanishing density, negative internal energy
    bad = np.array([[1e-20, 0.0, 0.0, 1.0], [1.0, 2.0, 0.0, 0.5]])
    q = _conv(L, 1, bad)
    assert q[0, 0] == 1e-10 and q[1, 3] == 1e-10
    # characteristic transform at reference states (1-D: rho, m, E)
    ref = p.copy()
    ch = _conv(L, 2, c, ref, width=3)
    back = _conv(L, 3, ch, ref)
    np.testing.assert_allclose(back[:, [0, 1, 3]], c[:, [0, 1, 3]],
                               rtol=1e-11, atol=1e-11)
    g = 1.4
    for k in range(10):
        rho, u, _, pr = ref[k]
        a = np.sqrt(g * pr / rho)
        H = a * a / (g - 1) + 0.5 * u * u
        A = np.array([[0.0, 1.0, 0.0],
                      [0.5 * (g - 3) * u * u, (3 - g) * u, g - 1],
                      [u * (0.5 * (g - 1) * u * u - H), H - (g - 1) * u * u,
                       g * u]])
        for j, lam in enumerate((u - a, u, u + a)):
            e = np.zeros((1, 3))
            e[0, j] = 1.0
            r = _conv(L, 3, e, ref[k:k + 1])[0, [0, 1, 3]]
            np.testing.assert_allclose(A @ r, lam * r, rtol=1e-12,
                                       atol=1e-12 * np.abs(r).max())


def test_limiter_functions():
    """the per-side limiter value of the device code (src/reconstruction.c:
    25-192) against the published functions: Barth-Jespersen
    min(1, bound / d); Venkatakrishnan, with epsilon = 0,
    (y^2 + 2y) / (y^2 + y + 2) for y = bound / d, tending to 1 as epsilon
    grows; the bound of the side d points to, and 1 where d = 0"""
    L = _lib()
    rng = np.random.default_rng(5)
    n = 400
    d = rng.normal(size=n)
    dmax = np.abs(rng.normal(size=n))
    dmin = -np.abs(rng.normal(size=n))
    d[:5] = 0.0

    def phi(lim, eps=0.0, d=d, dmax=dmax, dmin=dmin):
        out = np.empty(len(d))
        args = [np.ascontiguousarray(a) for a in (d, dmax, dmin)]
        L.phys_limiter(lim, len(d), *[a.ctypes.data_as(dp) for a in args],
                       C.c_double(eps), out.ctypes.data_as(dp))
        return out

    nz = d != 0
    y = np.where(d > 0, dmax, dmin)[nz] / d[nz]
    bj, vk = phi(1), phi(2)
    np.testing.assert_allclose(bj[nz], np.minimum(1.0, y), rtol=1e-15)
    np.testing.assert_allclose(vk[nz], (y * y + 2 * y) / (y * y + y + 2),
                               rtol=1e-13)
    assert np.all(bj[~nz] == 1.0) and np.all(vk[~nz] == 1.0)
    assert np.all((vk[nz] >= 0) & (vk[nz] < 1.0 + 1e-12) | (y > 1))
    assert abs(phi(2, d=np.array([1.0]), dmax=np.array([1.0]),
                   dmin=np.array([-1.0]))[0] - 0.75) < 1e-15
    # sign symmetry: the mirrored side gives the same value
    np.testing.assert_allclose(phi(2, d=-d, dmax=-dmin, dmin=-dmax), vk,
                               rtol=1e-15)
    # a large epsilon switches the limiter off
    assert np.all(np.abs(phi(2, eps=1e12)[nz] - 1.0) < 1e-9)


def test_viscous_flux_is_the_navier_stokes_flux():
    """diffusionFlux (src/fluxCalculation.c:855-873) of the device code
    against the textbook Navier-Stokes flux: stress tensor
    mu (grad u + grad u^T) - 2/3 mu (div u) I, its work u . tau, and
    Fourier conduction mu gamma / ((gamma - 1) Pr) grad(p / rho), the
    temperature gradient taken by central differences; a rigid rotation
    carries no stress"""
    L = _lib()
    rng = np.random.default_rng(2)
    gam, mu, Pr = 1.4, 3e-3, 0.72

    def visc(st, gx, gy):
        f, g = np.zeros(4), np.zeros(4)
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (st, gx, gy)]
        L.phys_visc(C.c_double(gam), C.c_double(mu), C.c_double(Pr),
                    *[x.ctypes.data_as(dp) for x in a], f.ctypes.data_as(dp),
                    g.ctypes.data_as(dp))
        return f, g

    k = mu * gam / ((gam - 1) * Pr)
    for _ in range(20):
        st = np.array([rng.uniform(0.5, 2), rng.normal(), rng.normal(),
                       rng.uniform(0.5, 2)])
        gx, gy = rng.normal(size=4), rng.normal(size=4)
        J = np.array([[gx[1], gy[1]], [gx[2], gy[2]]])   # d(u, v)/d(x, y)
        tau = mu * (J + J.T) - 2.0 / 3.0 * mu * np.trace(J) * np.eye(2)
        h = 1e-6
        for axis, (grad, flux) in enumerate(((gx, 0), (gy, 1))):
            T = [(st[3] + s * h * grad[3]) / (st[0] + s * h * grad[0])
                 for s in (1.0, -1.0)]
            dT = (T[0] - T[1]) / (2 * h)
            want = np.array([0.0, tau[0, axis], tau[1, axis],
                             st[1] * tau[0, axis] + st[2] * tau[1, axis] +
                             k * dT])
            got = visc(st, gx, gy)[flux]
            np.testing.assert_allclose(got, want, rtol=1e-8, atol=1e-12)
    # rigid rotation u = -w y, v = w x: no stress, no work
    w = 0.7
    st = np.array([1.0, 0.3, -0.2, 1.0])
    f, g = visc(st, np.array([0, 0, w, 0.0]), np.array([0, -w, 0, 0.0]))
    assert np.all(np.abs(f) < 1e-18) and np.all(np.abs(g) < 1e-18)


@pytest.mark.parametrize("code", list(range(1, 12)))
def test_fluxes_are_consistent(code):
    """every numerical flux of the device code is consistent: with equal
    states on both sides it returns the Euler flux of that state in the
    face-normal frame, (rho u, rho u^2 + p, rho u v, u (E + p))"""
    L = _lib()
    rng = np.random.default_rng(code)
    n = 200
    s = np.column_stack([rng.uniform(0.2, 2, n), rng.uniform(-2, 2, n),
                         rng.uniform(-2, 2, n), rng.uniform(0.2, 2, n)])
    s = np.ascontiguousarray(s)
    F = np.empty_like(s)
    L.phys_riemann(code, 0, C.c_double(1.4), n, s.ctypes.data_as(dp),
                   s.ctypes.data_as(dp), F.ctypes.data_as(dp))
    rho, u, v, p = s.T
    E = p / 0.4 + 0.5 * rho * (u * u + v * v)
    want = np.column_stack([rho * u, rho * u * u + p, rho * u * v,
                            u * (E + p)])
    np.testing.assert_allclose(F, want, rtol=1e-10, atol=1e-12)
