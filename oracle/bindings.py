"""ctypes bindings of the oracle (oracle/libccfd_oracle.so) and of the
reference harness (oracle/_ref/libccfdref*.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, never by the product package.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)


def _d(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(dp)


class OracleCfg(C.Structure):
    """struct oracle_cfg (oracle/ccfd_oracle.h)"""
    _fields_ = [
        ("gam", C.c_double), ("mu", C.c_double), ("Pr", C.c_double),
        ("navierstokes", C.c_int), ("iFlux", C.c_int),
        ("spatialOrder", C.c_int), ("limiter", C.c_int),
        ("venk_k", C.c_double), ("refState", (C.c_double * 4) * 2),
        ("rp1Dinterface", C.c_double), ("dxRef", C.c_double),
        ("cfl", C.c_double), ("dfl", C.c_double), ("timeStep1D", C.c_int),
        ("nRKstages", C.c_int), ("timeOrder", C.c_int),
        ("sourceFunc", C.c_int),
    ]


class ImplicitCfg(C.Structure):
    _fields_ = [("nKdim", C.c_int), ("nNewtonIter", C.c_int),
                ("precond", C.c_int), ("epsNewton", C.c_double),
                ("epsGMRES", C.c_double), ("gammaEW", C.c_double)]


def make_cfg(gamma=1.4, mu=0.0, Pr=0.72, navierstokes=0, flux=5, order=2,
             limiter=1, venk_k=1.0, left=(1, 0, 0, 1), right=(0.125, 0, 0, 0.1),
             x0=0.5, dxref=0.0, cfl=0.9, dfl=0.9, ts1d=0, nrk=1, torder=1,
             source=0):
    c = OracleCfg()
    c.gam, c.mu, c.Pr = gamma, mu, Pr
    c.navierstokes, c.iFlux, c.spatialOrder = navierstokes, flux, order
    c.limiter, c.venk_k = limiter, venk_k
    for i in range(4):
        c.refState[0][i] = left[i]
        c.refState[1][i] = right[i]
    c.rp1Dinterface, c.dxRef = x0, dxref
    c.cfl, c.dfl, c.timeStep1D = cfl, dfl, ts1d
    c.nRKstages, c.timeOrder = nrk, torder
    c.sourceFunc = source
    return c


_olib = None


def oracle_lib():
    global _olib
    if _olib is None:
        path = os.path.join(HERE, "libccfd_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: make -C oracle")
        L = C.CDLL(path)
        L.oracle_create.restype = C.c_void_p
        L.oracle_create.argtypes = [C.c_void_p, C.POINTER(OracleCfg)]
        L.oracle_destroy.argtypes = [C.c_void_p]
        L.oracle_set_prim.argtypes = [C.c_void_p, dp]
        L.oracle_get_prim.argtypes = [C.c_void_p, dp]
        L.oracle_get_cons.argtypes = [C.c_void_p, dp]
        L.oracle_fvtd.argtypes = [C.c_void_p, C.c_double, dp, dp, dp, dp, dp]
        L.oracle_riemann.argtypes = [C.POINTER(OracleCfg), C.c_int, dp, dp, dp]
        L.oracle_calc_dt.restype = C.c_double
        L.oracle_calc_dt.argtypes = [C.c_void_p, C.c_double, C.c_double,
                                     C.c_double, ip, ip]
        L.oracle_explicit_step.argtypes = [C.c_void_p, C.c_double, C.c_double, dp]
        L.oracle_global_residual.argtypes = [C.c_void_p, dp, dp]
        L.oracle_dot.restype = C.c_double
        L.oracle_dot.argtypes = [C.c_long, dp, dp]
        L.oracle_implicit_step.restype = C.c_int
        L.oracle_implicit_step.argtypes = [
            C.c_void_p, C.POINTER(ImplicitCfg), C.c_double, C.c_double, dp,
            ip, ip]
        L.oracle_build_matrix.restype = C.c_int
        L.oracle_build_matrix.argtypes = [C.c_void_p, C.c_double, C.c_double,
                                          dp, dp, dp]
        L.oracle_lusgs.restype = C.c_int
        L.oracle_lusgs.argtypes = [C.c_void_p, dp, dp, dp, dp]
        L.oracle_gmres.restype = C.c_int
        L.oracle_gmres.argtypes = [C.c_void_p, C.POINTER(ImplicitCfg),
                                   C.c_double, C.c_double, dp, dp, ip]
        L.oracle_rk_coeffs.restype = C.c_int
        L.oracle_rk_coeffs.argtypes = [C.c_int, C.c_int, dp]
        lp = C.POINTER(C.c_long)
        L.oracle_quadrature.argtypes = [C.c_void_p, C.c_long, dp, dp]
        L.oracle_quadrature.restype = C.c_int
        L.oracle_calc_errors.argtypes = [C.c_void_p, C.c_int, C.c_double,
                                         dp, dp, dp]
        L.oracle_wing_lists.argtypes = [C.c_void_p, C.c_int, lp, lp, lp]
        L.oracle_wing_lists.restype = C.c_long
        L.oracle_wing_coef.argtypes = [C.c_void_p, lp, C.c_long, C.c_long,
                                       C.c_double, dp, C.c_double, dp, dp,
                                       dp, dp]
        L.oracle_locate.argtypes = [C.c_void_p, dp]
        L.oracle_locate.restype = C.c_long
        L.oracle_exact_func.argtypes = [C.POINTER(OracleCfg), C.c_void_p,
                                        C.c_int, dp, C.c_double, dp]
        _olib = L
    return _olib


class Oracle:
    """CPU restatement of the operator and time loop on one mesh."""

    def __init__(self, mesh, cfg):
        self.L = oracle_lib()
        self.mesh = mesh
        self.cfg = cfg
        self.p = self.L.oracle_create(C.cast(mesh.ptr, C.c_void_p), C.byref(cfg))
        if not self.p:
            raise RuntimeError("oracle_create failed")
        self.N = mesh.n_elems
        self.S = mesh.n_slots

    def __del__(self):
        if getattr(self, "p", None):
            self.L.oracle_destroy(self.p)
            self.p = None

    def set_prim(self, prim):
        prim = np.ascontiguousarray(prim, dtype=np.float64)
        self.L.oracle_set_prim(self.p, _d(prim))

    def prim(self):
        out = np.empty((self.N, 4))
        self.L.oracle_get_prim(self.p, _d(out))
        return out

    def cons(self):
        out = np.empty((self.N, 4))
        self.L.oracle_get_cons(self.p, _d(out))
        return out

    def fvtd(self, time=0.0, full=False):
        ut = np.empty((self.N, 4))
        if not full:
            self.L.oracle_fvtd(self.p, time, _d(ut), None, None, None, None)
            return ut
        ux = np.empty((self.N, 4))
        uy = np.empty((self.N, 4))
        sp = np.empty((self.S, 4))
        fl = np.empty((self.S, 4))
        self.L.oracle_fvtd(self.p, time, _d(ut), _d(ux), _d(uy), _d(sp), _d(fl))
        return dict(u_t=ut, u_x=ux, u_y=uy, side=sp, flux=fl)

    def calc_dt(self, t, ptime, tend):
        vd, nf = C.c_int(), C.c_int()
        dt = self.L.oracle_calc_dt(self.p, t, ptime, tend, C.byref(vd), C.byref(nf))
        return dt, bool(vd.value), bool(nf.value)

    def explicit_step(self, t, dt):
        r = np.zeros(4)
        self.L.oracle_explicit_step(self.p, t, dt, _d(r))
        return r

    def implicit_step(self, ic, t, dt):
        r = np.zeros(4)
        nn, ng = C.c_int(0), C.c_int(0)
        rc = self.L.oracle_implicit_step(self.p, C.byref(ic), t, dt, _d(r),
                                         C.byref(nn), C.byref(ng))
        return rc, r, nn.value, ng.value

    def build_matrix(self, time, dt):
        """blocks of I - dt dR/dpVar at the current state: diagonal Dg
        (N, 4, 4), per slot Blk (S, 4, 4) (row: the slot's element, column:
        its neighbour) and the inverses Dinv (N, 4, 4) of the diagonal ones"""
        Dg = np.zeros((self.N, 4, 4))
        Blk = np.zeros((self.S, 4, 4))
        Dinv = np.zeros((self.N, 4, 4))
        rc = self.L.oracle_build_matrix(self.p, time, dt, _d(Dg), _d(Blk),
                                        _d(Dinv))
        if rc:
            raise RuntimeError("oracle_build_matrix failed (%d)" % rc)
        return Dg, Blk, Dinv

    def lusgs(self, Dinv, Blk, B):
        """one LU-SGS application to B (N, 4) with the blocks above"""
        Dinv = np.ascontiguousarray(Dinv, dtype=np.float64)
        Blk = np.ascontiguousarray(Blk, dtype=np.float64)
        B = np.ascontiguousarray(B, dtype=np.float64)
        X = np.zeros((self.N, 4))
        rc = self.L.oracle_lusgs(self.p, _d(Dinv), _d(Blk), _d(B), _d(X))
        if rc:
            raise RuntimeError("oracle_lusgs failed (%d)" % rc)
        return X

    def gmres(self, ic, t, dt, B):
        """one GMRES_M solve of (I - dt dR/dU) dX = -B at the current
        state; returns (dX (N, 4), Krylov iterations)"""
        B = np.ascontiguousarray(B, dtype=np.float64)
        dX = np.zeros((self.N, 4))
        n = C.c_int(0)
        rc = self.L.oracle_gmres(self.p, C.byref(ic), t, dt, _d(B), _d(dX),
                                 C.byref(n))
        if rc:
            raise RuntimeError("oracle_gmres failed (%d)" % rc)
        return dX, n.value

    def calc_errors(self, func, time):
        """calcErrors: (L1, L2, Linf), 4 values each"""
        L1, L2, Li = np.zeros(4), np.zeros(4), np.zeros(4)
        self.L.oracle_calc_errors(self.p, func, time, _d(L1), _d(L2), _d(Li))
        return L1, L2, Li

    def wing_lists(self, wall_code):
        """initWing: (slots, n_pressure, n_suction)"""
        slots = np.zeros(max(self.mesh.n_ghosts, 1), dtype=np.int64)
        nP, nS = C.c_long(), C.c_long()
        n = self.L.oracle_wing_lists(self.p, wall_code,
                                     slots.ctypes.data_as(C.POINTER(C.c_long)),
                                     C.byref(nP), C.byref(nS))
        return slots[:n].copy(), nP.value, nS.value

    def wing_coef(self, slots, n_pressure, n_suction, ref_length, ref, alpha):
        """calcCoef on the side states of the last fvtd: (cl, cd, cp, xyphi)"""
        slots = np.ascontiguousarray(slots, dtype=np.int64)
        n = len(slots)
        cp, xyphi = np.zeros(max(n, 1)), np.zeros((max(n, 1), 3))
        cl, cd = C.c_double(), C.c_double()
        ref = np.asarray(ref, dtype=np.float64)
        self.L.oracle_wing_coef(self.p, slots.ctypes.data_as(C.POINTER(C.c_long)),
                                n_pressure, n_suction, ref_length, _d(ref),
                                alpha, C.byref(cl), C.byref(cd), _d(cp),
                                _d(xyphi))
        return cl.value, cd.value, cp[:n], xyphi[:n]

    def residual(self, ut):
        ut = np.ascontiguousarray(ut, dtype=np.float64)
        r = np.zeros(4)
        self.L.oracle_global_residual(self.p, _d(ut), _d(r))
        return r


def exact_func(cfg, mesh, f, x, time):
    """exact function f at point x and time (src/exactFunction.c)"""
    out = np.zeros(4)
    x = np.ascontiguousarray(x, dtype=np.float64)
    oracle_lib().oracle_exact_func(C.byref(cfg), C.cast(mesh.ptr, C.c_void_p),
                                   f, _d(x), time, _d(out))
    return out


def quadrature(mesh, e):
    """quadrature points (n, 2) and weights (n) of element e"""
    x, w = np.zeros((5, 2)), np.zeros(5)
    n = oracle_lib().oracle_quadrature(C.cast(mesh.ptr, C.c_void_p), e, _d(x), _d(w))
    return x[:n], w[:n]


def locate(mesh, x):
    """initRecordPoints' element search: element index or -1"""
    x = np.asarray(x, dtype=np.float64)
    return oracle_lib().oracle_locate(C.cast(mesh.ptr, C.c_void_p), _d(x))


def dot(a, b):
    """vectorDotProduct (src/linearSolver.c:115-128) of two N x 4 vectors,
    summed serially in element order"""
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return oracle_lib().oracle_dot(len(a), _d(a), _d(b))


def riemann(cfg, code, L, R):
    lib = oracle_lib()
    L = np.ascontiguousarray(L, dtype=np.float64).reshape(-1, 4)
    R = np.ascontiguousarray(R, dtype=np.float64).reshape(-1, 4)
    F = np.empty_like(L)
    for i in range(len(L)):
        lib.oracle_riemann(C.byref(cfg), code, _d(L[i]), _d(R[i]), _d(F[i]))
    return F


def rk_coeffs(nrk, order):
    c = np.zeros(6)
    rc = oracle_lib().oracle_rk_coeffs(nrk, order, _d(c))
    return c if rc == 0 else None


# ------------------------------------------------------------ reference harness

def ref_available(ns=False):
    name = "libccfdref_ns.so" if ns else "libccfdref.so"
    return os.path.exists(os.path.join(HERE, "_ref", name))


class Ref:
    """The reference's own fvTimeDerivative / convectiveFlux via the harness.
    One instance per process and library (the reference keeps globals)."""

    _cache = {}

    def __init__(self, ns=False, omp=False, v3=False):
        name = "libccfdref_ns.so" if ns else (
            ("libccfdref_omp_v3.so" if v3 else "libccfdref_omp.so") if omp
            else "libccfdref.so")
        path = os.path.join(HERE, "_ref", name)
        if path not in Ref._cache:
            if not os.path.exists(path):
                raise RuntimeError("reference harness not built: " + path)
            L = C.CDLL(path, mode=C.RTLD_LOCAL)
            L.ref_init.argtypes = [C.c_char_p]
            L.ref_set_mesh.argtypes = [C.c_void_p]
            L.ref_set_mesh.restype = C.c_int
            L.ref_set_prim.argtypes = [dp]
            L.ref_fvtd.argtypes = [C.c_double, dp, dp, dp, dp, dp, dp]
            L.ref_riemann.argtypes = [C.c_int, C.c_long, dp, dp, dp]
            L.ref_set_riemann.argtypes = [dp, dp, C.c_double]
            L.ref_set_constants.argtypes = [C.c_double, C.c_double, C.c_double]
            L.ref_set_flux.argtypes = [C.c_int]
            Ref._cache[path] = L
        self.L = Ref._cache[path]

    def setup(self, ini_path, mesh, left=None, right=None, x0=0.5):
        self.L.ref_free()
        self.L.ref_init(ini_path.encode())
        if left is not None:
            self.L.ref_set_riemann(_d(np.asarray(left, dtype=np.float64)),
                                   _d(np.asarray(right, dtype=np.float64)), x0)
        rc = self.L.ref_set_mesh(C.cast(mesh.ptr, C.c_void_p))
        if rc != 0:
            raise RuntimeError("ref_set_mesh failed: %d" % rc)
        self.N = mesh.n_elems
        self.S = mesh.n_slots
        self.G = mesh.n_ghosts

    def set_prim(self, prim):
        prim = np.ascontiguousarray(prim, dtype=np.float64)
        self.L.ref_set_prim(_d(prim))

    def fvtd(self, time=0.0):
        ut = np.empty((self.N, 4))
        ux = np.empty((self.N, 4))
        uy = np.empty((self.N, 4))
        sp = np.empty((self.S, 4))
        fl = np.empty((self.S, 4))
        gp = np.empty((max(self.G, 1), 4))
        self.L.ref_fvtd(time, _d(ut), _d(ux), _d(uy), _d(sp), _d(fl), _d(gp))
        return dict(u_t=ut, u_x=ux, u_y=uy, side=sp, flux=fl, ghost=gp[:self.G])

    def set_flux(self, code):
        """iFlux of later fvtd calls"""
        self.L.ref_set_flux(code)

    def fvtd_ut(self, time=0.0):
        """dU/dt only"""
        ut = np.empty((self.N, 4))
        self.L.ref_fvtd(time, _d(ut), None, None, None, None, None)
        return ut

    def fvtd_only(self, time=0.0):
        self.L.ref_fvtd(time, None, None, None, None, None, None)

    def riemann(self, code, L, R, gamma=1.4):
        self.L.ref_set_constants(gamma, 0.0, 0.72)
        L = np.ascontiguousarray(L, dtype=np.float64).reshape(-1, 4)
        R = np.ascontiguousarray(R, dtype=np.float64).reshape(-1, 4)
        F = np.empty_like(L)
        self.L.ref_riemann(code, len(L), _d(L), _d(R), _d(F))
        return F


def write_ini(path, items):
    """Write a reference-style ini: items is a list of (key, value) pairs;
    repeated keys are positional, as the reference's parser consumes them."""
    with open(path, "w") as f:
        for k, v in items:
            if isinstance(v, (list, tuple)):
                v = "(/" + ", ".join(repr(float(x)) for x in v) + "/)"
            f.write("%s = %s\n" % (k, v))
