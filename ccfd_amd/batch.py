"""Python front of include/ccfd_batch.h: many explicit Euler solves in one
device arena, one launch per stage over all of them.

`Batch` admits instances (mesh + state + time-step settings), steps them
together, and reports per-instance time, step count and residuals. Errors
raise `GpuError`; `add` raises `NoRoom` when the arena is full.
"""
import ctypes as C

import numpy as np

from . import _native as N
from .gpu import GpuError, Params, lib as _gpu_lib


class NoRoom(GpuError):
    pass


class BatchCfg(C.Structure):
    _fields_ = [("scheme", Params), ("maxSides", C.c_int),
                ("localTimeStep", C.c_int), ("hbmBytes", C.c_size_t),
                ("maxInstances", C.c_int), ("maxHalo", C.c_int)]


class BatchCase(C.Structure):
    _fields_ = [("cfl", C.c_double), ("pTime", C.c_double),
                ("stopTime", C.c_double)]


class BatchInfo(C.Structure):
    _fields_ = [("nElems", C.c_long), ("steps", C.c_long), ("t", C.c_double),
                ("dt", C.c_double), ("res", C.c_double * 4),
                ("res0", C.c_double * 4), ("nonfinite", C.c_int)]


_blib = None
ENOMEM = -2


def lib():
    global _blib
    if _blib is not None:
        return _blib
    L = _gpu_lib()
    vp, dp, ip = C.c_void_p, N.c_double_p, N.c_int_p
    sig = {
        "ccfd_batch_create": (C.c_int, [C.POINTER(vp), C.c_int,
                                        C.POINTER(BatchCfg)]),
        "ccfd_batch_destroy": (C.c_int, [vp]),
        "ccfd_batch_bytes": (C.c_size_t, [vp, vp]),
        "ccfd_batch_fits": (C.c_int, [vp, vp, ip]),
        "ccfd_batch_add": (C.c_int, [vp, vp, C.POINTER(BatchCase), dp, ip]),
        "ccfd_batch_add_raw": (C.c_int, [vp, vp, C.POINTER(BatchCase), dp,
                                         ip]),
        "ccfd_batch_fits_raw": (C.c_int, [vp, vp, ip]),
        "ccfd_batch_remove": (C.c_int, [vp, C.c_int]),
        "ccfd_batch_step": (C.c_int, [vp, C.c_int]),
        "ccfd_batch_sync": (C.c_int, [vp]),
        "ccfd_batch_info_get": (C.c_int, [vp, C.c_int, C.POINTER(BatchInfo)]),
        "ccfd_batch_poll": (C.c_int, [vp, C.c_int, ip, C.POINTER(BatchInfo),
                                      ip]),
        "ccfd_batch_get_prim": (C.c_int, [vp, C.c_int, dp]),
        "ccfd_batch_get_cons": (C.c_int, [vp, C.c_int, dp]),
        "ccfd_batch_set_variant": (C.c_int, [vp, C.c_int]),
        "ccfd_batch_get_variant": (C.c_int, [vp, ip]),
        "ccfd_batch_set_graph": (C.c_int, [vp, C.c_int]),
        "ccfd_batch_autotune": (C.c_int, [vp, C.c_int, ip, dp]),
        "ccfd_batch_usage": (C.c_int, [vp, C.POINTER(C.c_long),
                                       C.POINTER(C.c_long), ip]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _blib = L
    return L


def _check(rc, L):
    if rc != 0:
        msg = L.ccfd_gpu_last_error()
        text = "ccfd_batch error %d: %s" % (rc, msg.decode() if msg else "")
        raise (NoRoom if rc == ENOMEM else GpuError)(text)


def info_dict(i):
    return {"n": i.nElems, "steps": i.steps, "t": i.t, "dt": i.dt,
            "res": np.array(i.res[:]), "res0": np.array(i.res0[:]),
            "nonfinite": bool(i.nonfinite)}


class Batch:
    """One arena on one device."""

    def __init__(self, scheme, max_sides=3, local_dt=False, hbm_bytes=0,
                 max_instances=0, max_halo=0, device=0):
        self.L = lib()
        cfg = BatchCfg()
        cfg.scheme = scheme
        cfg.maxSides, cfg.localTimeStep = max_sides, int(local_dt)
        cfg.hbmBytes, cfg.maxInstances, cfg.maxHalo = (int(hbm_bytes),
                                                       max_instances, max_halo)
        h = C.c_void_p()
        _check(self.L.ccfd_batch_create(C.byref(h), device, C.byref(cfg)),
               self.L)
        self.h = h
        self.n_elems = {}

    def close(self):
        if getattr(self, "h", None):
            self.L.ccfd_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def bytes_for(self, mesh):
        return int(self.L.ccfd_batch_bytes(self.h, C.cast(mesh.ptr, C.c_void_p)))

    def fits(self, mesh):
        f = C.c_int()
        _check(self.L.ccfd_batch_fits(self.h, C.cast(mesh.ptr, C.c_void_p),
                                      C.byref(f)), self.L)
        return bool(f.value)

    def add(self, mesh, prim, cfl=0.9, ptime=1e300, stop_time=1e300):
        prim = np.ascontiguousarray(prim, dtype=np.float64)
        assert prim.shape == (mesh.n_elems, 4)
        cs = BatchCase(cfl, ptime, stop_time)
        k = C.c_int()
        _check(self.L.ccfd_batch_add(self.h, C.cast(mesh.ptr, C.c_void_p),
                                     C.byref(cs), N.dptr(prim), C.byref(k)),
               self.L)
        self.n_elems[k.value] = mesh.n_elems
        return k.value

    def fits_raw(self, raw_mesh):
        f = C.c_int()
        _check(self.L.ccfd_batch_fits_raw(
            self.h, C.cast(raw_mesh.ptr, C.c_void_p), C.byref(f)), self.L)
        return bool(f.value)

    def add_raw(self, raw_mesh, prim, cfl=0.9, ptime=1e300, stop_time=1e300):
        """admit a mesh.RawMesh, preprocessed on the device"""
        prim = np.ascontiguousarray(prim, dtype=np.float64)
        assert prim.shape == (raw_mesh.n_elems, 4)
        cs = BatchCase(cfl, ptime, stop_time)
        k = C.c_int()
        _check(self.L.ccfd_batch_add_raw(
            self.h, C.cast(raw_mesh.ptr, C.c_void_p), C.byref(cs),
            N.dptr(prim), C.byref(k)), self.L)
        self.n_elems[k.value] = raw_mesh.n_elems
        return k.value

    def remove(self, k):
        _check(self.L.ccfd_batch_remove(self.h, k), self.L)
        self.n_elems.pop(k, None)

    def step(self, n):
        _check(self.L.ccfd_batch_step(self.h, n), self.L)

    def sync(self):
        _check(self.L.ccfd_batch_sync(self.h), self.L)

    def info(self, k):
        i = BatchInfo()
        _check(self.L.ccfd_batch_info_get(self.h, k, C.byref(i)), self.L)
        return info_dict(i)

    def poll(self, max_ids=4096):
        ids = np.zeros(max_ids, dtype=np.int32)
        infos = (BatchInfo * max_ids)()
        n = C.c_int()
        _check(self.L.ccfd_batch_poll(self.h, max_ids, N.iptr(ids), infos,
                                      C.byref(n)), self.L)
        return {int(ids[j]): info_dict(infos[j]) for j in range(n.value)}

    def _get(self, fn, k):
        out = np.empty((self.n_elems[k], 4))
        _check(fn(self.h, k, N.dptr(out)), self.L)
        return out

    def prim(self, k):
        return self._get(self.L.ccfd_batch_get_prim, k)

    def cons(self, k):
        return self._get(self.L.ccfd_batch_get_cons, k)

    def set_variant(self, v):
        _check(self.L.ccfd_batch_set_variant(self.h, v), self.L)

    def variant(self):
        v = C.c_int()
        _check(self.L.ccfd_batch_get_variant(self.h, C.byref(v)), self.L)
        return v.value

    def set_graph(self, on):
        _check(self.L.ccfd_batch_set_graph(self.h, int(on)), self.L)

    def autotune(self, nsteps=10):
        v = C.c_int()
        ms = np.zeros(4)
        _check(self.L.ccfd_batch_autotune(self.h, nsteps, C.byref(v),
                                          N.dptr(ms)), self.L)
        return v.value, ms

    def usage(self):
        u, c, n = C.c_long(), C.c_long(), C.c_int()
        _check(self.L.ccfd_batch_usage(self.h, C.byref(u), C.byref(c),
                                       C.byref(n)), self.L)
        return u.value, c.value, n.value


def riemann4_state(bary, seed, gamma=1.4, split=(0.5, 0.5)):
    """4-quadrant 2-D Riemann problem split at `split`: per quadrant
    rho, p ~ U[0.1, 2], u, v ~ U[-1, 1] (PCG64 with the given seed)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    q = np.empty((4, 4))
    q[:, 0] = rng.uniform(0.1, 2.0, 4)
    q[:, 3] = rng.uniform(0.1, 2.0, 4)
    q[:, 1] = rng.uniform(-1.0, 1.0, 4)
    q[:, 2] = rng.uniform(-1.0, 1.0, 4)
    x, y = bary[:, 0], bary[:, 1]
    quad = (x >= split[0]).astype(int) + 2 * (y >= split[1]).astype(int)
    return q[quad].copy()
