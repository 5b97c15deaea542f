"""Python front of the C ABI in include/ccfd_gpu.h (libccfd_gpu.so).

`Solver` owns one device context: upload a `Mesh` and parameters, set the
state, evaluate fvTimeDerivative, step in time. Errors raise `GpuError`
with the library's message; there is no CPU fallback.
"""
import ctypes as C

import numpy as np

from . import _native as N

FP64, FP32_FLUX = 0, 1
# operator implementations (enum ccfd_kernel_variant)
KV_AUTO, KV_SPLIT, KV_FUSED, KV_FUSED_FACE = 0, 1, 2, 3


class GpuError(RuntimeError):
    pass


class Params(C.Structure):
    """struct ccfd_gpu_params"""
    _fields_ = [
        ("gam", C.c_double), ("mu", C.c_double), ("Pr", C.c_double),
        ("navierstokes", C.c_int), ("iFlux", C.c_int),
        ("spatialOrder", C.c_int), ("limiter", C.c_int),
        ("venk_k", C.c_double), ("refState", (C.c_double * 4) * 2),
        ("rp1Dinterface", C.c_double), ("cfl", C.c_double),
        ("dfl", C.c_double), ("timeStep1D", C.c_int),
        ("nRKstages", C.c_int), ("timeOrder", C.c_int),
        ("precision", C.c_int),
    ]


class ImplicitCfg(C.Structure):
    """struct ccfd_gpu_implicit_cfg"""
    _fields_ = [("nKdim", C.c_int), ("nNewtonIter", C.c_int),
                ("precond", C.c_int), ("epsNewton", C.c_double),
                ("epsGMRES", C.c_double), ("gammaEW", C.c_double),
                ("jacobianOrder", C.c_int), ("sweeps", C.c_int),
                ("cycles", C.c_int)]


def params(gamma=1.4, mu=0.0, Pr=0.72, navierstokes=0, flux=5, order=2,
           limiter=1, venk_k=1.0, left=(1, 0, 0, 1), right=(0.125, 0, 0, 0.1),
           x0=0.5, cfl=0.9, dfl=0.9, ts1d=0, nrk=1, torder=1,
           precision=FP64):
    p = Params()
    p.gam, p.mu, p.Pr, p.navierstokes = gamma, mu, Pr, navierstokes
    p.iFlux, p.spatialOrder, p.limiter, p.venk_k = flux, order, limiter, venk_k
    for i in range(4):
        p.refState[0][i] = left[i]
        p.refState[1][i] = right[i]
    p.rp1Dinterface, p.cfl, p.dfl, p.timeStep1D = x0, cfl, dfl, ts1d
    p.nRKstages, p.timeOrder, p.precision = nrk, torder, precision
    return p


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    L = N._load("libccfd_gpu.so")
    vp, dp, ip = C.c_void_p, N.c_double_p, N.c_int_p
    sig = {
        "ccfd_gpu_last_error": (C.c_char_p, []),
        "ccfd_gpu_create": (C.c_int, [C.POINTER(vp), C.c_int]),
        "ccfd_gpu_destroy": (C.c_int, [vp]),
        "ccfd_gpu_set_mesh": (C.c_int, [vp, C.c_void_p, C.POINTER(Params)]),
        "ccfd_gpu_set_params": (C.c_int, [vp, C.POINTER(Params)]),
        "ccfd_gpu_set_prim": (C.c_int, [vp, dp]),
        "ccfd_gpu_set_cons": (C.c_int, [vp, dp]),
        "ccfd_gpu_get_prim": (C.c_int, [vp, dp]),
        "ccfd_gpu_get_cons": (C.c_int, [vp, dp]),
        "ccfd_gpu_get_ut": (C.c_int, [vp, dp]),
        "ccfd_gpu_get_recon": (C.c_int, [vp, dp, dp, dp]),
        "ccfd_gpu_fvtd": (C.c_int, [vp, C.c_double]),
        "ccfd_gpu_calc_dt": (C.c_int, [vp, C.c_double, C.c_double, C.c_double,
                                       dp, ip]),
        "ccfd_gpu_explicit_step": (C.c_int, [vp, C.c_double, C.c_double, dp]),
        "ccfd_gpu_advance": (C.c_int, [vp, C.c_int, dp, C.c_double,
                                       C.c_double, dp]),
        "ccfd_gpu_residual": (C.c_int, [vp, dp]),
        "ccfd_gpu_implicit_step": (C.c_int, [vp, C.POINTER(ImplicitCfg),
                                             C.c_double, C.c_double, dp, ip,
                                             ip]),
        "ccfd_gpu_riemann_batch": (C.c_int, [C.c_int, C.c_int, C.c_double,
                                             C.c_long, dp, dp, dp]),
        "ccfd_gpu_synchronize": (C.c_int, [vp]),
        "ccfd_gpu_set_variant": (C.c_int, [vp, C.c_int]),
        "ccfd_gpu_implicit_steady": (C.c_int, [vp, C.POINTER(ImplicitCfg),
                                               C.c_double, C.c_int,
                                               C.c_double, dp, dp, ip, ip,
                                               ip]),
        "ccfd_gpu_get_variant": (C.c_int, [vp, C.POINTER(C.c_int)]),
        "ccfd_gpu_set_tile": (C.c_int, [vp, C.c_int]),
        "ccfd_gpu_get_tile": (C.c_int, [vp, C.POINTER(C.c_int)]),
        "ccfd_gpu_autotune": (C.c_int, [vp, C.c_int, C.POINTER(C.c_int),
                                        C.POINTER(C.c_int), dp]),
        "ccfd_gpu_profile_steps": (C.c_int, [vp, C.c_int, C.c_double,
                                             C.c_double, C.c_int, dp, ip,
                                             dp]),
        "ccfd_gpu_copy_probe": (C.c_int, [C.c_int, C.c_size_t, C.c_int, dp]),
        "ccfd_gpu_gmres": (C.c_int, [vp, C.POINTER(ImplicitCfg), C.c_double,
                                     C.c_double, C.c_double, C.c_double, dp,
                                     C.c_double, dp, ip]),
        "ccfd_gpu_dot": (C.c_int, [vp, dp, dp, dp]),
        "ccfd_gpu_jacobian_blocks": (C.c_int, [vp, C.POINTER(ImplicitCfg),
                                               C.c_double, C.c_double,
                                               C.c_int, dp, dp]),
        "ccfd_gpu_set_source": (C.c_int, [vp, C.c_int]),
        "ccfd_gpu_get_bc_sides": (C.c_int, [vp, dp]),
        "ccfd_gpu_wing_init": (C.c_int, [vp, C.c_void_p, C.c_int, C.c_double,
                                         C.POINTER(C.c_long),
                                         C.POINTER(C.c_long), dp]),
        "ccfd_gpu_wing_coef": (C.c_int, [vp, dp, C.c_double, dp, dp]),
        "ccfd_gpu_calc_errors": (C.c_int, [vp, C.c_int, C.c_double, dp, dp,
                                           dp]),
        "ccfd_gpu_get_prim_at": (C.c_int, [vp, C.c_int, C.POINTER(C.c_long),
                                           dp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc):
    if rc != 0:
        msg = lib().ccfd_gpu_last_error()
        raise GpuError("ccfd_gpu error %d: %s" % (rc, msg.decode() if msg else ""))


def riemann_batch(code, L, R, gamma=1.4, precision=FP64):
    """convectiveFlux on n rotated state pairs, evaluated on the GPU."""
    L = np.ascontiguousarray(L, dtype=np.float64).reshape(-1, 4)
    R = np.ascontiguousarray(R, dtype=np.float64).reshape(-1, 4)
    F = np.empty_like(L)
    check(lib().ccfd_gpu_riemann_batch(code, precision, gamma, len(L),
                                       N.dptr(L), N.dptr(R), N.dptr(F)))
    return F


def copy_probe(nbytes=1 << 30, reps=20, device=0):
    """streaming-copy bandwidth (GB/s, read + write) on the device"""
    g = C.c_double()
    check(lib().ccfd_gpu_copy_probe(device, nbytes, reps, C.byref(g)))
    return g.value


class Solver:
    """One device context with a resident mesh and state."""

    def __init__(self, mesh, prm, device=0):
        self.L = lib()
        h = C.c_void_p()
        check(self.L.ccfd_gpu_create(C.byref(h), device))
        self.h = h
        self.mesh = mesh
        self.prm = prm
        check(self.L.ccfd_gpu_set_mesh(self.h, C.cast(mesh.ptr, C.c_void_p),
                                       C.byref(prm)))
        self.N = mesh.n_elems
        self.S = mesh.n_slots

    def set_variant(self, variant):
        """Operator implementation: KV_AUTO, KV_SPLIT, KV_FUSED or
        KV_FUSED_FACE."""
        check(self.L.ccfd_gpu_set_variant(self.h, variant))

    def variant(self):
        """The implementation the next operator call runs."""
        v = C.c_int()
        check(self.L.ccfd_gpu_get_variant(self.h, C.byref(v)))
        return v.value

    def set_tile(self, elems):
        """elements per block of the face-once kernel: 0 (automatic), 256
        or 512"""
        check(self.L.ccfd_gpu_set_tile(self.h, elems))

    def autotune(self, reps=5):
        """time the operator implementations and tiles on the resident
        state and keep the fastest: (variant, tile, {name: ms per call})"""
        v, t = C.c_int(), C.c_int()
        ms = np.zeros(4)
        check(self.L.ccfd_gpu_autotune(self.h, reps, C.byref(v), C.byref(t),
                                       N.dptr(ms)))
        names = ("split", "fused", "fused_face_256", "fused_face_512")
        return v.value, t.value, {n: (m if m >= 0 else None)
                                  for n, m in zip(names, ms.tolist())}

    def tile(self):
        """the tile the next operator call runs"""
        v = C.c_int()
        check(self.L.ccfd_gpu_get_tile(self.h, C.byref(v)))
        return v.value

    def close(self):
        if getattr(self, "h", None):
            self.L.ccfd_gpu_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_params(self, prm):
        self.prm = prm
        check(self.L.ccfd_gpu_set_params(self.h, C.byref(prm)))

    def set_prim(self, prim):
        prim = np.ascontiguousarray(prim, dtype=np.float64)
        assert prim.shape == (self.N, 4)
        check(self.L.ccfd_gpu_set_prim(self.h, N.dptr(prim)))

    def set_cons(self, cons):
        cons = np.ascontiguousarray(cons, dtype=np.float64)
        check(self.L.ccfd_gpu_set_cons(self.h, N.dptr(cons)))

    def _get(self, fn):
        out = np.empty((self.N, 4))
        check(fn(self.h, N.dptr(out)))
        return out

    def prim(self):
        return self._get(self.L.ccfd_gpu_get_prim)

    def cons(self):
        return self._get(self.L.ccfd_gpu_get_cons)

    def ut(self):
        return self._get(self.L.ccfd_gpu_get_ut)

    def recon(self):
        ux = np.empty((self.N, 4))
        uy = np.empty((self.N, 4))
        sp = np.empty((self.S, 4))
        check(self.L.ccfd_gpu_get_recon(self.h, N.dptr(ux), N.dptr(uy),
                                        N.dptr(sp)))
        return ux, uy, sp

    def fvtd(self, time=0.0):
        check(self.L.ccfd_gpu_fvtd(self.h, time))

    # ---- source term and analysis consumers (src/source.c, src/analyze.c)
    def set_source(self, func):
        """calcSource with sourceFunction func (0: off)"""
        check(self.L.ccfd_gpu_set_source(self.h, func))

    def bc_sides(self):
        """interior side states at the boundary sides (ghost order) from
        the last operator evaluation"""
        out = np.empty((max(self.mesh.n_ghosts, 1), 4))
        check(self.L.ccfd_gpu_get_bc_sides(self.h, N.dptr(out)))
        return out[:self.mesh.n_ghosts]

    def wing_init(self, wall_code, ref_length=1.0):
        """initWing: (n_pressure, n_suction, xyphi)"""
        nP, nS = C.c_long(), C.c_long()
        xy = np.zeros((max(self.mesh.n_ghosts, 1), 3))
        check(self.L.ccfd_gpu_wing_init(self.h, C.cast(self.mesh.ptr, C.c_void_p),
                                        wall_code, ref_length, C.byref(nP),
                                        C.byref(nS), N.dptr(xy)))
        return nP.value, nS.value, xy[:nP.value + nS.value]

    def wing_coef(self, ref, alpha, n_sides):
        """calcCoef: (cl, cd, cp)"""
        ref = np.ascontiguousarray(ref, dtype=np.float64)
        coef = np.zeros(2)
        cp = np.zeros(max(n_sides, 1))
        check(self.L.ccfd_gpu_wing_coef(self.h, N.dptr(ref), alpha,
                                        N.dptr(coef), N.dptr(cp)))
        return coef[0], coef[1], cp[:n_sides]

    def calc_errors(self, func, time):
        """calcErrors: (L1, L2, Linf)"""
        L1, L2, Li = np.zeros(4), np.zeros(4), np.zeros(4)
        check(self.L.ccfd_gpu_calc_errors(self.h, func, time, N.dptr(L1),
                                          N.dptr(L2), N.dptr(Li)))
        return L1, L2, Li

    def prim_at(self, elems):
        """primitive state of the given elements (reference numbering)"""
        e = np.ascontiguousarray(elems, dtype=np.int64)
        out = np.zeros((max(len(e), 1), 4))
        check(self.L.ccfd_gpu_get_prim_at(self.h, len(e),
                                          e.ctypes.data_as(C.POINTER(C.c_long)),
                                          N.dptr(out)))
        return out[:len(e)]

    def calc_dt(self, t, ptime, tend):
        dt = C.c_double()
        vd = C.c_int()
        check(self.L.ccfd_gpu_calc_dt(self.h, t, ptime, tend, C.byref(dt),
                                      C.byref(vd)))
        return dt.value, bool(vd.value)

    def explicit_step(self, t, dt, residual=True):
        r = np.zeros(4)
        check(self.L.ccfd_gpu_explicit_step(self.h, t, dt,
                                            N.dptr(r) if residual else None))
        return r

    def advance(self, nsteps, t, ptime, tend, residual=True):
        tt = C.c_double(t)
        r = np.zeros(4)
        check(self.L.ccfd_gpu_advance(self.h, nsteps, C.byref(tt), ptime, tend,
                                      N.dptr(r) if residual else None))
        return tt.value, r

    def residual(self):
        r = np.zeros(4)
        check(self.L.ccfd_gpu_residual(self.h, N.dptr(r)))
        return r

    def implicit_step(self, cfg, t, dt):
        r = np.zeros(4)
        nn, ng = C.c_int(0), C.c_int(0)
        check(self.L.ccfd_gpu_implicit_step(self.h, C.byref(cfg), t, dt,
                                            N.dptr(r), C.byref(nn),
                                            C.byref(ng)))
        return r, nn.value, ng.value

    def gmres(self, cfg, t_global, time, dt, alpha, B, normB=None):
        """GMRES_M about the resident state for the right-hand side B (N x 4,
        reference order); returns (dX, Krylov iterations)"""
        B = np.ascontiguousarray(B, dtype=np.float64)
        assert B.shape == (self.N, 4)
        if normB is None:
            normB = float(np.sqrt(self.dot(B, B)))
        dX = np.empty((self.N, 4))
        it = C.c_int(0)
        check(self.L.ccfd_gpu_gmres(self.h, C.byref(cfg), t_global, time, dt,
                                    alpha, N.dptr(B), normB, N.dptr(dX),
                                    C.byref(it)))
        return dX, it.value

    def dot(self, a, b):
        """vectorDotProduct of two N x 4 host vectors, summed on the device"""
        a = np.ascontiguousarray(a, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        assert a.shape == b.shape == (self.N, 4)
        out = C.c_double(0.0)
        check(self.L.ccfd_gpu_dot(self.h, N.dptr(a), N.dptr(b), C.byref(out)))
        return out.value

    def jacobian_blocks(self, cfg, time, dt, method=1):
        """the preconditioner's blocks at the resident state (reference
        order): D (N, 4, 4) and the side blocks (nSlots, 4, 4); method 0
        coloured finite differences, 1 element by element"""
        D = np.empty((self.N, 4, 4))
        Bk = np.empty((self.S, 4, 4))
        check(self.L.ccfd_gpu_jacobian_blocks(self.h, C.byref(cfg), time, dt,
                                              method, N.dptr(D), N.dptr(Bk)))
        return D, Bk

    def implicit_steady(self, cfg, cfl, max_steps, drop, res0=None):
        """implicit pseudo-time steps with local steps until the density
        residual drops by `drop`; returns (res, res0, steps, newton, gmres).
        Passing the res0 of an earlier call continues that solve."""
        r = np.zeros(4)
        r0 = np.zeros(4) if res0 is None else np.array(res0, dtype=np.float64)
        st, nn, ng = C.c_int(0), C.c_int(0), C.c_int(0)
        check(self.L.ccfd_gpu_implicit_steady(
            self.h, C.byref(cfg), cfl, max_steps, drop, N.dptr(r), N.dptr(r0),
            C.byref(st), C.byref(nn), C.byref(ng)))
        return r, r0, st.value, nn.value, ng.value

    def sync(self):
        check(self.L.ccfd_gpu_synchronize(self.h))

    def profile_steps(self, nsteps, ptime=1e300, tend=1e300, residual=True):
        """average kernel durations (ms) over nsteps device-driven steps
        (each with globalResidual unless residual is False): dict(dt, recon,
        flux, other, step), launch counts and the last step's residual"""
        ms = np.zeros(5)
        cnt = (C.c_int * 4)()
        r = np.zeros(4)
        check(self.L.ccfd_gpu_profile_steps(self.h, nsteps, ptime, tend,
                                            1 if residual else 0, N.dptr(ms),
                                            cnt, N.dptr(r)))
        keys = ("dt", "recon", "flux", "other", "step")
        return dict(zip(keys, ms.tolist())), list(cnt), r
