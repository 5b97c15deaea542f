"""Served steady solves: requests arriving over time, solved with
continuous batching over `ccfd_amd.batch.Batch` arenas.

The serving loop itself (FIFO admission within the arenas' room, step
bursts, residual polling, retirement, precision promotion) is the C ABI of
include/ccfd_serve.h (ccfd_amd/csrc/serve.cpp); this module feeds it.
Requests arrive at given offsets (e.g. a Poisson stream); a thread pool
prepares each one's mesh and initial state as it arrives (mesh generation
and preprocessing run in C / numpy and release the GIL); prepared requests
are submitted in arrival order, and each round of the server admits what
fits, steps every admitted solve together and retires finished ones. Work
of requests that arrive while others run overlaps on the device: admission
happens between step bursts, not after the batch drains.

A solve retires when its density residual has dropped by the target factor
(relative to its first step), when it reaches the step cap, or when its
time step turns non-finite.

Precision promotion (optional `polish` arena): FP32 flux arithmetic floors
the residual near 1e-5 of its initial value, above a 1e-6 target. A solve
therefore runs in the main (FP32-flux) arena until its drop reaches
`switch_drop` (or stops improving within 10x of it), and then continues
from its conservative state (consPrim) in the polish (FP64) arena to the
target.

Arenas are either `Batch` objects (the server drives them from C) or any
object with the same methods (fits, add, remove, step, poll, cons, prim),
reached through callbacks: the CPU tests use stand-ins.

Reported per request: arrival, admission and completion times, steps (and
how many ran in the polish arena), residual drop, and whether it converged.
`summary` gives solves/s and latency percentiles.
"""
import collections
import concurrent.futures as cf
import ctypes as C
import time

import numpy as np

from . import _native as N
from . import batch as BT

CCFD_ENOMEM = -2

_CLOCK = C.CFUNCTYPE(C.c_double, C.c_void_p)
_FITS = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, C.POINTER(C.c_int))
_ADD = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, C.POINTER(C.c_double),
                   C.c_double, C.POINTER(C.c_int))
_REMOVE = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int)
_STEP = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int)
_POLL = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.POINTER(C.c_int),
                    C.POINTER(BT.BatchInfo), C.POINTER(C.c_int))
_GET = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.POINTER(C.c_double))


class Backend(C.Structure):
    """struct ccfd_serve_backend"""
    _fields_ = [("arena", C.c_void_p), ("fits", _FITS), ("add", _ADD),
                ("remove", _REMOVE), ("step", _STEP), ("poll", _POLL),
                ("get_cons", _GET), ("get_prim", _GET)]


class ServeCfg(C.Structure):
    """struct ccfd_serve_cfg"""
    _fields_ = [("targetDrop", C.c_double), ("maxSteps", C.c_long),
                ("stepsPerRound", C.c_int), ("maxActive", C.c_int),
                ("switchDrop", C.c_double), ("stallRounds", C.c_int),
                ("gamma", C.c_double), ("keepState", C.c_int),
                ("clock", _CLOCK), ("clockArg", C.c_void_p)]


class ServeResult(C.Structure):
    """struct ccfd_serve_result"""
    _fields_ = [("id", C.c_long), ("nElems", C.c_long), ("steps", C.c_long),
                ("polishSteps", C.c_long), ("drop", C.c_double),
                ("converged", C.c_int), ("nonfinite", C.c_int),
                ("admitted", C.c_double), ("done", C.c_double)]


def _lib():
    L = BT.lib()
    if getattr(L, "_serve_typed", False):
        return L
    vp = C.c_void_p
    L.ccfd_serve_backend_batch.argtypes = [vp, C.POINTER(Backend)]
    L.ccfd_serve_backend_batch_raw.argtypes = [vp, C.POINTER(Backend)]
    L.ccfd_serve_create.argtypes = [C.POINTER(vp), C.POINTER(Backend),
                                    C.POINTER(Backend), C.POINTER(ServeCfg)]
    L.ccfd_serve_destroy.argtypes = [vp]
    L.ccfd_serve_submit.argtypes = [vp, C.c_long, vp, C.c_long,
                                    C.POINTER(C.c_double), C.c_double]
    L.ccfd_serve_round.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                   C.POINTER(C.c_int)]
    L.ccfd_serve_results.argtypes = [vp, C.c_int, C.POINTER(ServeResult),
                                     C.POINTER(C.c_int)]
    L.ccfd_serve_take_state.argtypes = [vp, C.c_long, C.POINTER(C.c_double)]
    for f in ("backend_batch", "create", "destroy", "submit", "round",
              "results", "take_state"):
        getattr(L, "ccfd_serve_" + f).restype = C.c_int
    L._serve_typed = True
    return L


class Request:
    """One solve: `make()` returns (mesh, prim, cfl); `meta` is free-form
    (kind, size, flow conditions) and is copied into the result."""

    def __init__(self, rid, arrival, make, meta=None):
        self.id = rid
        self.arrival = float(arrival)
        self.make = make
        self.meta = dict(meta or {})


def cons_to_prim(cons, gamma=1.4):
    """consPrim (src/equationOfState.c:37-52) on host arrays"""
    p = np.empty_like(cons)
    p[:, 0] = cons[:, 0]
    p[:, 1] = cons[:, 1] / cons[:, 0]
    p[:, 2] = cons[:, 2] / cons[:, 0]
    p[:, 3] = (gamma - 1.0) * (cons[:, 3] - 0.5 * (cons[:, 1] * p[:, 1] +
                                                    cons[:, 2] * p[:, 2]))
    p[:, 0] = np.where(p[:, 0] < 1e-19, 1e-10, p[:, 0])
    p[:, 3] = np.where(p[:, 3] < 1e-19, 1e-10, p[:, 3])
    return p


class _PyArena:
    """callbacks of a Python arena object (a stand-in in tests); meshes are
    referred to by integer handles the server passes back"""

    def __init__(self, obj, objs):
        self.obj, self.objs = obj, objs

        def fits(_, h, f):
            f[0] = int(bool(obj.fits(objs[h][0])))
            return 0

        def add(_, h, prim, cfl, k):
            # the submitted state object, or (on a hand-over to the polish
            # arena) the state the server converted
            mesh, state = objs[h]
            if prim:
                state = np.ctypeslib.as_array(
                    prim, shape=(mesh.n_elems, 4)).copy()
            k[0] = int(obj.add(mesh, state, cfl))
            return 0

        def remove(_, k):
            obj.remove(k)
            return 0

        def step(_, n):
            obj.step(n)
            return 0

        def poll(_, cap, ids, info, n):
            j = 0
            for k, inf in obj.poll().items():
                if j >= cap:
                    break
                ids[j] = k
                i = info[j]
                i.nElems = int(inf.get("n", 0))
                i.steps = int(inf["steps"])
                i.res[:] = [float(x) for x in inf["res"][:4]]
                i.res0[:] = [float(x) for x in inf["res0"][:4]]
                i.nonfinite = int(bool(inf["nonfinite"]))
                j += 1
            n[0] = j
            return 0

        def getter(name):
            def get(_, k, out):
                a = getattr(obj, name)(k)
                if a is not None:
                    a = np.asarray(a, dtype=np.float64).ravel()
                    C.memmove(out, a.ctypes.data, 8 * len(a))
                return 0
            return get

        self.fns = (_FITS(fits), _ADD(add), _REMOVE(remove), _STEP(step),
                    _POLL(poll), _GET(getter("cons")), _GET(getter("prim")))
        self.be = Backend(None, *self.fns)


class Server:
    def __init__(self, batch, target_drop=1e-6, max_steps=20000,
                 steps_per_poll=50, prep_workers=8, keep_state=False,
                 max_active=None, polish=None, switch_drop=1e-4,
                 stall_polls=4, gamma=1.4, clock=time.perf_counter,
                 sleep=time.sleep, log=None, log_every=10.0, raw=False,
                 autotune_steps=0):
        # raw: requests carry mesh.RawMesh objects, which a device arena
        # preprocesses on the GPU at admission (a stand-in arena gets the
        # object as it is)
        self.raw = raw
        # autotune_steps > 0: after a round that admitted solves to a device
        # arena, time its operator implementations on the new mix
        # (ccfd_batch_autotune; the state is restored) and keep the fastest
        self.autotune_steps = autotune_steps
        self.tuned = []                # (arena, variant) per tuning
        self.batch = batch
        self.polish = polish           # FP64 arena for the final drop
        self.prep_workers = prep_workers
        self.keep_state = keep_state
        self.clock = clock
        self.sleep = sleep
        self.log = log                 # progress lines, every log_every s
        self.log_every = log_every
        self.L = _lib()
        self._objs = {}                # handle -> (mesh, prim) for Python arenas
        self._next = 0
        self._arenas = []
        self.native = isinstance(getattr(batch, "h", None), C.c_void_p)
        main = self._backend(batch)
        pol = self._backend(polish) if polish is not None else None
        self._clock = _CLOCK(lambda _: float(self.clock()))
        cfg = ServeCfg(target_drop, max_steps, steps_per_poll,
                       int(max_active or 0), switch_drop, stall_polls, gamma,
                       int(keep_state), self._clock, None)
        h = C.c_void_p()
        self._check(self.L.ccfd_serve_create(
            C.byref(h), C.byref(main), C.byref(pol) if pol else None,
            C.byref(cfg)))
        self.h = h

    def _backend(self, arena):
        # a device arena (ccfd_amd.batch.Batch) is driven from C; anything
        # else with the same methods through callbacks
        native = isinstance(getattr(arena, "h", None), C.c_void_p)
        if native != self.native:
            raise ValueError("both arenas must be device batches or both "
                             "stand-ins")
        if native:
            be = Backend()
            fn = (self.L.ccfd_serve_backend_batch_raw if self.raw
                  else self.L.ccfd_serve_backend_batch)
            self._check(fn(arena.h, C.byref(be)))
            return be
        a = _PyArena(arena, self._objs)
        self._arenas.append(a)
        return a.be

    def _check(self, rc):
        if rc != 0:
            msg = self.L.ccfd_gpu_last_error()
            msg = msg.decode() if msg else ""
            raise (ValueError if rc == CCFD_ENOMEM else RuntimeError)(msg)

    def close(self):
        if getattr(self, "h", None):
            self.L.ccfd_serve_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _submit(self, r, mesh, prim, cfl, inflight):
        if self.native:
            prim = np.ascontiguousarray(prim, dtype=np.float64)
            assert prim.shape == (mesh.n_elems, 4)
            handle = C.cast(mesh.ptr, C.c_void_p).value
            pp = N.dptr(prim)
        else:
            self._next += 1
            handle = self._next
            self._objs[handle] = (mesh, prim)
            pp = None
        rec = dict(r.meta)
        rec.update(id=r.id, n=mesh.n_elems, arrival=r.arrival)
        # the mesh and state stay referenced until the result is back
        inflight[r.id] = (rec, mesh, prim, handle)
        self._check(self.L.ccfd_serve_submit(self.h, r.id, handle,
                                             mesh.n_elems, pp, cfl))

    def _collect(self, inflight, t0, results):
        buf = (ServeResult * 64)()
        while True:
            n = C.c_int()
            self._check(self.L.ccfd_serve_results(self.h, 64, buf, C.byref(n)))
            for j in range(n.value):
                x = buf[j]
                rec, mesh, prim, handle = inflight.pop(x.id)
                rec.update(admitted=x.admitted - t0, done=x.done - t0,
                           steps=int(x.steps), polish_steps=int(x.polishSteps),
                           drop=float(x.drop), converged=bool(x.converged),
                           nonfinite=bool(x.nonfinite))
                rec["latency"] = rec["done"] - rec["arrival"]
                if self.keep_state:
                    out = np.empty((mesh.n_elems, 4))
                    self._check(self.L.ccfd_serve_take_state(
                        self.h, x.id, N.dptr(out)))
                    rec["prim"] = out
                self._objs.pop(handle, None)
                results.append(rec)
            if n.value < 64:
                return

    def run(self, requests):
        reqs = sorted(requests, key=lambda r: (r.arrival, r.id))
        pending = collections.deque(reqs)     # not yet arrived
        prepping = collections.deque()        # (request, future), FIFO
        inflight = {}                         # submitted, not yet retired
        results = []
        t0 = self.clock()
        next_log = self.log_every
        busy = False
        prev = {}                             # solves per arena, last round
        with cf.ThreadPoolExecutor(self.prep_workers) as pool:
            while pending or prepping or busy:
                now = self.clock() - t0
                while pending and pending[0].arrival <= now:
                    r = pending.popleft()
                    prepping.append((r, pool.submit(r.make)))
                # prepared requests join the server's queue in arrival order
                while prepping and prepping[0][1].done():
                    r, fut = prepping.popleft()
                    mesh, prim, cfl = fut.result()
                    self._submit(r, mesh, prim, cfl, inflight)
                act, pol, que = C.c_int(), C.c_int(), C.c_int()
                self._check(self.L.ccfd_serve_round(
                    self.h, C.byref(act), C.byref(pol), C.byref(que)))
                self._collect(inflight, t0, results)
                if self.native and self.autotune_steps > 0:
                    for name, arena, n in (("main", self.batch, act.value),
                                           ("polish", self.polish, pol.value)):
                        if arena is not None and n > prev.get(name, 0):
                            v, _ = arena.autotune(self.autotune_steps)
                            self.tuned.append((name, v))
                    prev = {"main": act.value, "polish": pol.value}
                busy = act.value + pol.value + que.value > 0
                if self.log and now >= next_log:
                    next_log = now + self.log_every
                    self.log("serve: %.1f s, %d done, %d active, %d polishing, "
                             "%d waiting" % (now, len(results), act.value,
                                             pol.value, que.value + len(prepping)
                                             + len(pending)))
                if act.value == 0 and pol.value == 0:
                    # idle: wait for the next arrival or preparation
                    wait = 0.002
                    if pending and not prepping:
                        wait = max(0.0, min(0.05, pending[0].arrival - now))
                    self.sleep(wait)
        return sorted(results, key=lambda x: x["id"])


def summary(results, makespan=None):
    """Solves/s and latency percentiles (seconds) of a run. A solve counts
    when it reached the target residual drop; requests retired at the step
    cap or with a non-finite state are failures, counted apart, and the
    retirement rate over all requests is given beside the solve rate."""
    if not results:
        return {"solves": 0, "requests": 0}
    conv = [r for r in results if r["converged"]]
    capped = [r for r in results if not r["converged"] and not r.get("nonfinite")]
    nonfin = [r for r in results if not r["converged"] and r.get("nonfinite")]
    t_end = max(r["done"] for r in results)
    t_beg = min(r["arrival"] for r in results)
    span = makespan if makespan is not None else t_end - t_beg

    def pct(rs, q):
        return float(np.percentile([r["latency"] for r in rs], q)) if rs else None
    return {
        "requests": len(results),
        "solves": len(conv),
        "failed_step_cap": len(capped),
        "failed_nonfinite": len(nonfin),
        "makespan_s": span,
        "solves_per_s": len(conv) / span if span > 0 else float("inf"),
        "retirements_per_s": len(results) / span if span > 0 else float("inf"),
        "latency_p50_s": pct(conv, 50),
        "latency_p99_s": pct(conv, 99),
        "latency_p50_all_s": pct(results, 50),
        "latency_p99_all_s": pct(results, 99),
        "latency_max_s": float(max(r["latency"] for r in results)),
        "cells_total": int(sum(r["n"] for r in results)),
        "steps_total": int(sum(r["steps"] for r in results)),
        "polish_steps_total": int(sum(r.get("polish_steps", 0)
                                      for r in results)),
    }


def poisson_arrivals(n, rate, seed):
    """n arrival offsets (s) of a Poisson stream with the given rate (1/s)"""
    rng = np.random.Generator(np.random.PCG64(seed))
    return np.cumsum(rng.exponential(1.0 / rate, n))
