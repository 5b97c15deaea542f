"""Served steady solves: a request loop with continuous batching over
`ccfd_amd.batch.Batch` arenas.

Requests arrive over time (the caller gives arrival offsets, e.g. a Poisson
stream). Each request names how to build its mesh and initial state; a
thread pool prepares them as they arrive (mesh generation and
preprocessing run in C / numpy and release the GIL). The loop admits
prepared requests in arrival order whenever the arena has room (HBM
budget), steps every admitted solve together, and retires a solve when its
density residual has dropped by the target factor (relative to its first
step), when it reaches the step cap, or when its time step turns
non-finite. Work of requests that arrive while others run overlaps on the
device: admission happens between step bursts, not after the batch
drains.

Precision promotion (optional `polish` arena): FP32 flux arithmetic floors
the residual near 1e-5 of its initial value, above a 1e-6 target. A solve
therefore runs in the main (FP32-flux) arena until its drop reaches
`switch_drop` (or stops improving within 10x of it), and then continues
from its FP64 conservative state in the polish (FP64) arena to the target. Both arenas
step concurrently on their own streams.

Reported per request: arrival, admission and completion times, steps (and
how many ran in the polish arena), residual drop, and whether it converged.
`summary` gives solves/s and latency percentiles.
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


def cons_to_prim(cons, gamma=1.4):
    """consPrim (src/equationOfState.c:37-52) on host arrays"""
    p = np.empty_like(cons)
    rho = np.where(cons[:, 0] < 1e-19, 1e-10, cons[:, 0])
    p[:, 0] = rho
    p[:, 1] = cons[:, 1] / rho
    p[:, 2] = cons[:, 2] / rho
    pr = (gamma - 1.0) * (cons[:, 3] - 0.5 * rho * (p[:, 1] ** 2 + p[:, 2] ** 2))
    p[:, 3] = np.where(pr < 1e-19, 1e-10, pr)
    return p


class Server:
    def __init__(self, batch, target_drop=1e-6, max_steps=20000,
                 steps_per_poll=50, prep_workers=8, keep_state=False,
                 max_active=None, polish=None, switch_drop=1e-4,
                 stall_polls=4, gamma=1.4, clock=time.perf_counter,
                 sleep=time.sleep, log=None, log_every=10.0):
        self.batch = batch
        self.polish = polish           # FP64 arena for the final drop
        self.switch_drop = switch_drop
        self.stall_polls = stall_polls
        self.gamma = gamma
        self.max_active = max_active   # 1: one solve at a time
        self.target_drop = target_drop
        self.max_steps = max_steps
        self.steps_per_poll = steps_per_poll
        self.prep_workers = prep_workers
        self.keep_state = keep_state
        self.clock = clock
        self.sleep = sleep
        self.log = log                 # progress lines, every log_every s
        self.log_every = log_every

    def _finish(self, arena, k, rec, inf, drop, conv, t0, results):
        if self.keep_state:
            rec["prim"] = arena.prim(k)
        arena.remove(k)
        rec.update(done=self.clock() - t0, steps=rec["steps0"] + inf["steps"],
                   drop=drop, converged=bool(conv),
                   nonfinite=inf["nonfinite"])
        for key in ("steps0", "res0", "hist", "mesh", "state", "cfl"):
            rec.pop(key, None)
        rec["latency"] = rec["done"] - rec["arrival"]
        results.append(rec)

    def run(self, requests):
        reqs = sorted(requests, key=lambda r: (r.arrival, r.id))
        pending = collections.deque(reqs)     # not yet arrived
        prepping = collections.deque()        # (request, future), FIFO
        active = {}                           # main arena id -> record
        polishing = {}                        # polish arena id -> record
        moving = collections.deque()          # records waiting for polish room
        results = []
        t0 = self.clock()
        next_log = self.log_every
        with cf.ThreadPoolExecutor(self.prep_workers) as pool:
            while pending or prepping or active or polishing or moving:
                now = self.clock() - t0
                if self.log and now >= next_log:
                    next_log = now + self.log_every
                    self.log("serve: %.1f s, %d done, %d active (%d cells), "
                             "%d polishing, %d waiting" % (
                                 now, len(results), len(active),
                                 sum(r["n"] for r in active.values()),
                                 len(polishing) + len(moving),
                                 len(prepping) + len(pending)))
                while pending and pending[0].arrival <= now:
                    r = pending.popleft()
                    prepping.append((r, pool.submit(r.make)))
                # solves handed over to the polish arena, in order
                if moving and not polishing and \
                        not self.polish.fits(moving[0]["mesh"]):
                    raise ValueError("request %d does not fit the empty "
                                     "polish arena" % moving[0]["id"])
                while moving and self.polish.fits(moving[0]["mesh"]):
                    rec = moving.popleft()
                    k = self.polish.add(rec["mesh"], rec.pop("state"),
                                        cfl=rec["cfl"])
                    polishing[k] = rec
                # admit prepared requests in arrival order while they fit
                while prepping and prepping[0][1].done() and (
                        self.max_active is None or
                        len(active) + len(polishing) + len(moving) <
                        self.max_active):
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
                               admitted=self.clock() - t0, cfl=cfl, steps0=0,
                               res0=None, hist=[], polish_steps=0,
                               mesh=mesh if self.polish else None)
                    active[k] = rec
                    del mesh, prim
                if not (active or polishing):
                    # idle: wait for the next arrival, preparation or room
                    wait = 0.002
                    if pending and not prepping and not moving:
                        wait = max(0.0, min(0.05, pending[0].arrival - now))
                    self.sleep(wait)
                    continue
                if active:
                    self.batch.step(self.steps_per_poll)
                if polishing:
                    self.polish.step(self.steps_per_poll)
                if active:
                    self._poll_main(active, moving, t0, results)
                if polishing:
                    self._poll_polish(polishing, t0, results)
        return sorted(results, key=lambda x: x["id"])

    def _poll_main(self, active, moving, t0, results):
        for k, inf in self.batch.poll().items():
            if k not in active:
                continue
            rec = active[k]
            if rec["res0"] is None:
                rec["res0"] = inf["res0"][0]
            drop = inf["res"][0] / rec["res0"] if rec["res0"] > 0 else 0.0
            conv = drop <= self.target_drop
            capped = inf["steps"] >= self.max_steps
            if conv or capped or inf["nonfinite"]:
                active.pop(k)
                self._finish(self.batch, k, rec, inf, drop, conv, t0, results)
                continue
            if self.polish is None:
                continue
            rec["hist"].append(drop)
            h = rec["hist"]
            # stalled near the FP32 floor; a solve that stalls far above it
            # (unsteady flow, limiter cycling) would not converge in FP64
            # either and stays on the cheaper arithmetic until the cap
            stalled = (len(h) > self.stall_polls and
                       drop <= 10.0 * self.switch_drop and
                       min(h[-self.stall_polls:]) >
                       0.5 * min(h[:-self.stall_polls]))
            if drop <= self.switch_drop or stalled:
                # continue from the FP64 conservative state in the polish arena
                rec["state"] = cons_to_prim(self.batch.cons(k), self.gamma)
                rec["steps0"] = inf["steps"]
                self.batch.remove(k)
                active.pop(k)
                moving.append(rec)

    def _poll_polish(self, polishing, t0, results):
        for k, inf in self.polish.poll().items():
            if k not in polishing:
                continue
            rec = polishing[k]
            drop = inf["res"][0] / rec["res0"] if rec["res0"] > 0 else 0.0
            conv = drop <= self.target_drop
            rec["polish_steps"] = inf["steps"]
            if conv or rec["steps0"] + inf["steps"] >= self.max_steps or \
                    inf["nonfinite"]:
                polishing.pop(k)
                self._finish(self.polish, k, rec, inf, drop, conv, t0, results)


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
