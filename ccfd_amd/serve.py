"""Served steady solves: a request loop with continuous batching over one
`ccfd_amd.batch.Batch`.

Requests arrive over time (the caller gives arrival offsets, e.g. a Poisson
stream). Each request names how to build its mesh and initial state; a
thread pool prepares them as they arrive (mesh generation and
preprocessing run in C / numpy and release the GIL). The loop admits
prepared requests in arrival order whenever the arena has room (HBM
budget), steps every admitted solve together, and retires a solve when its
density residual has dropped by the target factor, when it reaches the
step cap, or when its time step turns non-finite. Work of requests that
arrive while others run overlaps on the device: admission happens between
step bursts, not after the batch drains.

Reported per request: arrival, admission and completion times, steps,
residual drop, and whether it converged. `summary` gives solves/s and
latency percentiles.
"""
import collections
import concurrent.futures as cf
import time

import numpy as np


class Request:
    """One solve: `make()` returns (mesh, prim, cfl); `meta` is free-form
    (kind, size, flow conditions) and is copied into the result."""

    def __init__(self, rid, arrival, make, meta=None):
        self.id = rid
        self.arrival = float(arrival)
        self.make = make
        self.meta = dict(meta or {})


class Server:
    def __init__(self, batch, target_drop=1e-6, max_steps=20000,
                 steps_per_poll=50, prep_workers=8, keep_state=False,
                 max_active=None, clock=time.perf_counter, sleep=time.sleep):
        self.batch = batch
        self.max_active = max_active   # 1: one solve at a time
        self.target_drop = target_drop
        self.max_steps = max_steps
        self.steps_per_poll = steps_per_poll
        self.prep_workers = prep_workers
        self.keep_state = keep_state
        self.clock = clock
        self.sleep = sleep

    def run(self, requests):
        reqs = sorted(requests, key=lambda r: (r.arrival, r.id))
        pending = collections.deque(reqs)     # not yet arrived
        prepping = collections.deque()        # (request, future), FIFO
        active = {}                           # batch id -> (request, record)
        results = []
        t0 = self.clock()
        with cf.ThreadPoolExecutor(self.prep_workers) as pool:
            while pending or prepping or active:
                now = self.clock() - t0
                while pending and pending[0].arrival <= now:
                    r = pending.popleft()
                    prepping.append((r, pool.submit(r.make)))
                # admit prepared requests in arrival order while they fit
                while prepping and prepping[0][1].done() and (
                        self.max_active is None or
                        len(active) < self.max_active):
                    r, fut = prepping[0]
                    mesh, prim, cfl = fut.result()
                    if not self.batch.fits(mesh):
                        if not active:
                            raise ValueError(
                                "request %d (%d cells) does not fit the empty "
                                "arena" % (r.id, mesh.n_elems))
                        break
                    prepping.popleft()
                    k = self.batch.add(mesh, prim, cfl=cfl)
                    rec = dict(r.meta)
                    rec.update(id=r.id, n=mesh.n_elems, arrival=r.arrival,
                               admitted=self.clock() - t0)
                    active[k] = (r, rec)
                    del mesh, prim
                if active:
                    self.batch.step(self.steps_per_poll)
                    for k, inf in self.batch.poll().items():
                        if k not in active:
                            continue
                        drop = inf["res"][0] / inf["res0"][0] \
                            if inf["res0"][0] > 0 else 0.0
                        conv = drop <= self.target_drop
                        if conv or inf["steps"] >= self.max_steps or \
                                inf["nonfinite"]:
                            r, rec = active.pop(k)
                            if self.keep_state:
                                rec["prim"] = self.batch.prim(k)
                            self.batch.remove(k)
                            rec.update(done=self.clock() - t0,
                                       steps=inf["steps"], drop=drop,
                                       converged=bool(conv),
                                       nonfinite=inf["nonfinite"])
                            rec["latency"] = rec["done"] - rec["arrival"]
                            results.append(rec)
                else:
                    # idle: wait for the next arrival or preparation
                    wait = 0.002
                    if pending and not prepping:
                        wait = max(0.0, min(0.05, pending[0].arrival - now))
                    self.sleep(wait)
        return sorted(results, key=lambda x: x["id"])


def summary(results, makespan=None):
    """solves/s over the run and latency percentiles (seconds)"""
    if not results:
        return {"solves": 0}
    lat = np.array([r["latency"] for r in results])
    t_end = max(r["done"] for r in results)
    t_beg = min(r["arrival"] for r in results)
    span = makespan if makespan is not None else t_end - t_beg
    return {
        "solves": len(results),
        "converged": int(sum(r["converged"] for r in results)),
        "makespan_s": span,
        "solves_per_s": len(results) / span if span > 0 else float("inf"),
        "latency_p50_s": float(np.percentile(lat, 50)),
        "latency_p99_s": float(np.percentile(lat, 99)),
        "latency_max_s": float(lat.max()),
        "cells_total": int(sum(r["n"] for r in results)),
        "steps_total": int(sum(r["steps"] for r in results)),
    }


def poisson_arrivals(n, rate, seed):
    """n arrival offsets (s) of a Poisson stream with the given rate (1/s)"""
    rng = np.random.Generator(np.random.PCG64(seed))
    return np.cumsum(rng.exponential(1.0 / rate, n))
