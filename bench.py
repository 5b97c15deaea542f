#!/usr/bin/env python3
"""Benchmark: explicit RK stages of ccfd's finite-volume operator on one
MI355X.

Workload (default): Delaunay triangulation of a jittered 1415 x 1415 point
grid on the periodic unit square (3,998,792 triangles), smooth seeded state
(density / pressure = 1 + 0.1 x 4 sine modes, Mach 0.5 flow at a random
angle), HLLC flux, 2nd order with Venkatakrishnan limiting (K = 5),
nRKstages = 4 / timeOrder = 2, FP64. A step is the reference's explicit
step: calcTimeStep, four fvTimeDerivative + RK stage updates and
globalResidual (src/timeDiscretization.c:362-401). The metric counts
cell-stage updates.

Prints one JSON line with the metric, a roofline entry for the dominant
kernel (the fused operator + stage update, or the flux + gather + stage
update kernel of the split pair) from HIP-event kernel timings over the
timed region, the same measurement for the Roe, HLL and AUSMD fluxes and
for the FP32-flux mode, the dU/dt of the initial state compared with the
reference's own fvTimeDerivative (compiled from its sources) on the same
mesh and state for every flux, and a CPU baseline: that reference
fvTimeDerivative timed with OpenMP on this host. Beside it, short batched
and served samples of the many-mesh workloads.
"""
import argparse
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
# algorithmic HBM bytes per element of the flux / gather / stage-update
# kernel (FP64, triangles, 2nd order): the operator's flux phase reads
# (element state 32, gradient 64, GP 48, side ids 12, face normal + length
# 36, 1/area 8 = 200 B) plus the stage update (read U0 32, write U 32,
# write primitive state 32); the dU/dt round trip is fused away.
B_FLUX_FP64 = 296.0
B_FLUX_FP32 = 186.0   # same with FP32 operator data (FP64 U0 / U)
# whole stage (fvTimeDerivative + update), FP64 / FP32-flux, triangles + VK:
# the canonical algorithmic bytes of the two-phase dataflow (limited
# gradients materialised once), used for the fused kernel, which runs the
# whole stage in one launch
B_STAGE_FP64 = 572.0
B_STAGE_FP32 = 330.0
FLUX_NAMES = {2: "roe", 3: "hll", 5: "hllc", 9: "ausmd"}
# the device code the committed traffic figures were measured on
KERNEL_SOURCES = ("fused.hpp", "operator.hpp", "physics.hpp", "layout.hpp",
                  "op_flux.hip", "soa.cpp")


# what the fused kernel itself must move per element and stage (each array
# once, side states in LDS), FP64: reads state 32 + links 12 + LSQ weights
# 48 + GP 48 + eps^2 8 + side normals 48 + signed lengths 24 + 1/area 8 +
# U0 32 = 260, writes the primitive state 32; the last stage also writes U
# 32 and dU/dt 32. FP32 operator data: 152 + 16, last stage + 32 + 16.
# Halo re-reads come on top.
def fused_min_bytes(fp64, stages):
    per, last = (292.0, 64.0) if fp64 else (168.0, 48.0)
    return per + last / stages


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_case(n, seed):
    from ccfd_amd import mesh as M
    t0 = time.time()
    raw, bcs = M.jittered_delaunay(n, 0.3, seed, periodic=True)
    t1 = time.time()
    mesh = M.Mesh(raw, bcs)
    t2 = time.time()
    prim = M.smooth_state(mesh.bary, seed=seed)
    log("mesh: %d triangles, delaunay %.1f s, preprocessing %.1f s"
        % (mesh.n_elems, t1 - t0, t2 - t1))
    return raw, bcs, mesh, prim


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _has_x86_64_v3():
    """AVX2, FMA, BMI2 and MOVBE: what -march=x86-64-v3 code may use"""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    fl = set(line.split(":", 1)[1].split())
                    return {"avx2", "fma", "bmi2", "movbe", "f16c"} <= fl
    except OSError:
        pass
    return False


def _ref_ini(flux, venk_k):
    from oracle import bindings as B
    ini = "/tmp/ccfd_bench_ref_%d.ini" % os.getpid()
    B.write_ini(ini, [("fluxFunction", flux), ("nBC", 4),
                      ("BCtype", 701), ("connection", (0.0, 1.0)),
                      ("BCtype", 702), ("connection", (0.0, 0.0)),
                      ("BCtype", 711), ("connection", (1.0, 0.0)),
                      ("BCtype", 712), ("connection", (0.0, 0.0)),
                      ("spatialOrder", 2), ("limiter", 2),
                      ("venk_k", venk_k)])
    return ini


def reference_run(mesh, prim, flux, venk_k, seconds, parity_fluxes):
    """The reference's fvTimeDerivative, compiled from its own sources with
    OpenMP (oracle/_ref), on the same mesh and state: dU/dt of the state
    for every flux in parity_fluxes (for the parity check), then the main
    flux timed for about `seconds` (the CPU baseline). Returns
    (baseline dict or None, {flux: dU/dt})."""
    from oracle import bindings as B
    if not B.ref_available(False):
        return None, {}
    v3 = _has_x86_64_v3() and os.path.exists(
        os.path.join(ROOT, "oracle", "_ref", "libccfdref_omp_v3.so"))
    if not v3 and not os.path.exists(
            os.path.join(ROOT, "oracle", "_ref", "libccfdref_omp.so")):
        return None, {}
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(
        16, os.cpu_count() or 1)
    # read by libgomp when the library below loads it
    os.environ["OMP_NUM_THREADS"] = str(threads)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    ini = _ref_ini(flux, venk_k)
    devnull = os.open(os.devnull, os.O_WRONLY)
    saved = os.dup(1)
    os.dup2(devnull, 1)   # the reference prints its parameter echo
    uts = {}
    try:
        ref = B.Ref(omp=True, v3=v3)
        t0 = time.time()
        ref.setup(ini, mesh)
        t_setup = time.time() - t0
        ref.set_prim(prim)
        for code in parity_fluxes:
            ref.set_flux(code)
            uts[code] = ref.fvtd_ut(0.0)
        ref.set_flux(flux)
        ref.fvtd_only(0.0)   # warm-up
        calls, t0 = 0, time.time()
        while calls < 3 or time.time() - t0 < seconds:
            ref.fvtd_only(0.0)
            calls += 1
        el = time.time() - t0
    finally:
        # the reference's printf output sits in the C library's buffer:
        # flush it while stdout still points at /dev/null
        import ctypes
        ctypes.CDLL(None).fflush(None)
        os.dup2(saved, 1)
        os.close(devnull)
        os.close(saved)
    try:
        os.unlink(ini)
    except OSError:
        pass
    flags = ("-std=c99 -O3 -march=x86-64-v3 -flto -fopenmp" if v3 else
             "-std=c99 -O3 -fopenmp")
    base = {
        "value": mesh.n_elems * calls / el,
        "unit": "cell-evaluations/s (reference fvTimeDerivative, OpenMP)",
        "cores": threads,
        "kind": "reference",
        "sample": "%d fvTimeDerivative calls on the same %d-triangle mesh and "
                  "state (%.1f s; RK update, dt and residual, about 8%% of "
                  "the reference's CPU time per step, not included). Build: "
                  "the reference's own finiteVolume.c, reconstruction.c, "
                  "boundary.c, fluxCalculation.c, equationOfState.c and "
                  "helpers, gcc %s (the reference's Makefile flags with "
                  "-march=x86-64-v3 for -march=native). Data: the "
                  "reference's elem_t / side_t records and pointer links, "
                  "allocated as contiguous arrays by the harness "
                  "(oracle/ref_harness.c) instead of one malloc per record "
                  "by mesh.c (set-up %.1f s). OpenMP: %d threads, "
                  "OMP_PROC_BIND=%s, OMP_PLACES=%s. CPU: %s"
                  % (calls, mesh.n_elems, el, flags, t_setup, threads,
                     os.environ.get("OMP_PROC_BIND"),
                     os.environ.get("OMP_PLACES"), _cpu_model()),
    }
    return base, uts


def _reference_child(raw, bcs, prim, flux, venk_k, seconds, parity_fluxes):
    """reference_run in a process of its own (spawned): its OpenMP runtime,
    thread binding and host arrays stay out of the process that drives the
    GPU"""
    from ccfd_amd import mesh as M
    mesh = M.Mesh(raw, bcs)
    return reference_run(mesh, prim, flux, venk_k, seconds, parity_fluxes)


def reference_in_child(raw, bcs, prim, flux, venk_k, seconds, parity_fluxes):
    import concurrent.futures as cf
    import multiprocessing as mp
    with cf.ProcessPoolExecutor(1, mp_context=mp.get_context("spawn")) as ex:
        return ex.submit(_reference_child, raw, bcs, prim, flux, venk_k,
                         seconds, parity_fluxes).result()


def kernel_source_sha():
    """hash of the device sources the committed traffic figures depend on"""
    h = hashlib.sha256()
    for f in KERNEL_SOURCES:
        try:
            with open(os.path.join(ROOT, "ccfd_amd", "csrc", f), "rb") as fh:
                h.update(fh.read())
        except OSError:
            h.update(b"missing:" + f.encode())
    return h.hexdigest()[:16]


def pmc_traffic(kernel_key):
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 --pmc summary (FETCH_SIZE / WRITE_SIZE, copy-kernel scaled),
    with its source: None, and a note, when the device sources have changed
    since those counters were read."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None, "no committed counter summary"
    ent = d.get("kernels", {}).get(kernel_key)
    if ent is None:
        return None, "no counter figure for %s" % kernel_key
    sha = kernel_source_sha()
    if d.get("kernel_source_sha") != sha:
        return None, ("stale: profiles/pmc_traffic.json was measured on "
                      "device sources %s, these are %s"
                      % (d.get("kernel_source_sha"), sha))
    return ent["bytes_per_launch"], "profiles/%s (%s)" % (ent["summary"],
                                                        ent["command"])


def _ut_parity(got, ref, gamma=1.4):
    """dU/dt of the device against the reference's: bitwise, the largest
    difference relative to the largest value of its variable, and the
    relative L2 / Linf of the density and total-energy rates"""
    import numpy as np
    scale = np.maximum(np.abs(ref).max(axis=0), 1e-300)
    d = np.abs(got - ref)
    out = {"bitwise": bool(np.array_equal(got, ref)),
           "max_rel_err": float((d / scale).max())}
    for name, v in (("density", 0), ("energy", 3)):
        l2, linf = _rel_err(got[:, v], ref[:, v])
        out["%s_ut_rel_l2" % name] = l2
        out["%s_ut_rel_linf" % name] = linf
    return out


def _rel_err(a, b):
    """relative L2 and Linf of a against b"""
    import numpy as np
    d = a - b
    return (float(np.linalg.norm(d) / np.linalg.norm(b)),
            float(np.abs(d).max() / np.abs(b).max()))


def _density_err(a, b):
    """relative L2 and Linf of the density of primitive state a against b"""
    return _rel_err(a[:, 0], b[:, 0])


def _energy_err(a, b, gamma=1.4):
    """relative L2 and Linf of the total energy per volume, p / (gamma - 1)
    + rho |v|^2 / 2, of primitive state a against b"""
    def energy(q):
        return q[:, 3] / (gamma - 1.0) + 0.5 * q[:, 0] * (q[:, 1] ** 2 +
                                                         q[:, 2] ** 2)
    return _rel_err(energy(a), energy(b))


def _max_mach(prim, gamma=1.4):
    import numpy as np
    c = np.sqrt(gamma * prim[:, 3] / prim[:, 0])
    return float((np.hypot(prim[:, 1], prim[:, 2]) / c).max())


def _err_summary(errs, both=None):
    """max and median relative L2, max relative Linf over the solves with
    finite errors, and the count of the others; with a mask, the same
    maxima over the solves it marks"""
    import numpy as np

    ok = [e for e in errs if np.isfinite(e[0]) and np.isfinite(e[1])]
    out = {"l2_max": max([e[0] for e in ok] or [0.0]),
           "l2_median": float(np.median([e[0] for e in ok])) if ok else None,
           "linf_max": max([e[1] for e in ok] or [0.0]),
           # solves whose error is not finite (not in the maxima above)
           "nonfinite": len(errs) - len(ok)}
    if both is not None:
        sel = [e for e, m in zip(errs, both) if m]
        out.update(both_converged=len(sel),
                   both_converged_l2_max=max([e[0] for e in sel] or [0.0]),
                   both_converged_linf_max=max([e[1] for e in sel] or [0.0]))
    return out


def served_run(n_req, cells, rate, max_steps, hbm_gb, seed=5, baseline=True,
               venk_k=5.0, raw=True, variant=None):
    """Served steady solves: Poisson arrivals, continuous batching in an
    FP32-flux arena with local time steps, retire on a 1e-6 density
    residual drop (or the step cap); then the same requests one at a time in
    FP64. Returns the summary with p50 / p99 latency, solves/s of both and
    the density error of the FP32 results against the FP64 ones."""
    import numpy as np
    from ccfd_amd import batch as BT
    from ccfd_amd import gpu as G
    from ccfd_amd import serve as SV
    from ccfd_amd import workloads as W
    specs = W.served_specs(n_req, seed=seed, cells=cells)
    arr = SV.poisson_arrivals(n_req, rate, seed)

    def requests(arrivals):
        return [SV.Request(sp["id"], a,
                           (lambda sp=sp: W.served_case(sp, raw=raw)), sp)
                for sp, a in zip(specs, arrivals)]

    def scheme(prec):
        return G.params(flux=9, order=2, limiter=2, venk_k=venk_k, nrk=4,
                        torder=2, precision=prec)
    out = {"requests": n_req, "cells_range": list(cells),
           "arrival_rate_per_s": rate, "max_steps": max_steps,
           "target_drop": 1e-6,
           "scheme": "AUSMD, VK K=%g, RK4 local time steps at CFL 2; served: "
                     "FP32 flux / FP64 state until a 1e-4 drop or a stall, "
                     "then FP64 to the target; baseline: FP64 one solve at a "
                     "time; meshes %s" % (venk_k, "preprocessed on the device "
                     "at admission" if raw else "preprocessed on the host")}
    b = BT.Batch(scheme(G.FP32_FLUX), max_sides=4, local_dt=True,
                 hbm_bytes=(3 * hbm_gb << 30) // 4)
    pol = BT.Batch(scheme(G.FP64), max_sides=4, local_dt=True,
                   hbm_bytes=(hbm_gb << 30) // 4)
    if variant is not None:
        b.set_variant(variant)
        pol.set_variant(variant)
        out["variant"] = variant
    srv = SV.Server(b, target_drop=1e-6, max_steps=max_steps,
                    steps_per_poll=100, keep_state=baseline, polish=pol,
                    switch_drop=1e-4, log=log, raw=raw,
                    autotune_steps=0 if variant is not None else 4)
    res = srv.run(requests(arr))
    b.close()
    pol.close()
    out["served"] = SV.summary(res)
    names = {G.KV_AUTO: "auto", G.KV_SPLIT: "split", G.KV_FUSED: "fused",
             G.KV_FUSED_FACE: "fused_face"}

    def tuned(server):
        cnt = {}
        for arena, v in server.tuned:
            key = "%s_%s" % (arena, names.get(v, str(v)))
            cnt[key] = cnt.get(key, 0) + 1
        return cnt
    out["served"]["operator_choices"] = tuned(srv)
    log("served: %s" % json.dumps(out["served"]))
    if baseline:
        b64 = BT.Batch(scheme(G.FP64), max_sides=4, local_dt=True,
                       hbm_bytes=hbm_gb << 30)
        srv64 = SV.Server(b64, target_drop=1e-6, max_steps=max_steps,
                          steps_per_poll=100, keep_state=True, max_active=1,
                          raw=raw, autotune_steps=0 if variant is not None else 4,
                          log=log)
        res64 = srv64.run(requests(np.zeros(n_req)))
        one_choices = tuned(srv64)
        b64.close()
        out["one_at_a_time_fp64"] = SV.summary(res64)
        out["one_at_a_time_fp64"]["operator_choices"] = one_choices
        both = [r["converged"] and q["converged"] for r, q in zip(res, res64)]
        out["density_err_vs_fp64"] = _err_summary(
            [_density_err(r["prim"], q["prim"]) for r, q in zip(res, res64)],
            both)
        out["energy_err_vs_fp64"] = _err_summary(
            [_energy_err(r["prim"], q["prim"]) for r, q in zip(res, res64)],
            both)
        one = out["one_at_a_time_fp64"]
        # solves: requests that reached the target drop; capped and
        # non-finite retirements are failures in both columns
        out["speedup_solves_per_s"] = (
            out["served"]["solves_per_s"] / one["solves_per_s"]
            if one.get("solves_per_s") else None)
        out["converged_in_both"] = int(sum(both))
        # the requests both modes solved, each mode timed over its own run
        out["speedup_on_requests_both_solved"] = (
            (sum(both) / out["served"]["makespan_s"]) /
            (sum(both) / one["makespan_s"]) if sum(both) else None)
    return out


def served_convergence_run(n_req, cells, venk_ks, chunks, chunk, hbm_gb,
                           seed=5):
    """Residual history of the served requests without the serving loop:
    every request in one FP64 arena (the served scheme: AUSMD, VK, RK4 local
    steps at CFL 2, free-stream start), stepped in chunks for each
    Venkatakrishnan constant; per request the density residual drop after
    every chunk and the first chunk end at or below 1e-6."""
    from ccfd_amd import batch as BT
    from ccfd_amd import gpu as G
    from ccfd_amd import workloads as W
    specs = W.served_specs(n_req, seed=seed, cells=cells)
    out = {"requests": n_req, "cells_range": list(cells), "chunk": chunk,
           "steps": chunks * chunk, "target_drop": 1e-6,
           "scheme": "AUSMD, VK, RK4 local time steps at CFL 2, FP64, "
                     "meshes preprocessed on the device at admission",
           "runs": {}}
    for K in venk_ks:
        sch = G.params(flux=9, order=2, limiter=2, venk_k=K, nrk=4, torder=2,
                       precision=G.FP64)
        b = BT.Batch(sch, max_sides=4, local_dt=True, hbm_bytes=hbm_gb << 30)
        ids = {}
        for sp in specs:
            m, p, cfl = W.served_case(sp, raw=True)
            ids[b.add_raw(m, p, cfl=cfl)] = sp
        variant, _ = b.autotune(4)
        hist = {k: [] for k in ids}
        nonfinite = set()
        t0 = time.perf_counter()
        for c in range(chunks):
            b.step(chunk)
            b.sync()
            info = b.poll()
            for k in ids:
                i = info[k]
                hist[k].append(float(i["res"][0] / i["res0"][0])
                               if i["res0"][0] > 0 else float("nan"))
                if i["nonfinite"]:
                    nonfinite.add(k)
            log("K=%g: %d steps, %.1f s" % (K, (c + 1) * chunk,
                                             time.perf_counter() - t0))
        el = time.perf_counter() - t0
        reqs = []
        for k in sorted(ids, key=lambda k: ids[k]["id"]):
            sp, h = ids[k], hist[k]
            first = next((j for j, d in enumerate(h) if d <= 1e-6), None)
            reqs.append({"id": sp["id"], "kind": sp["kind"],
                         "cells": sp["cells"], "mach": round(sp["mach"], 3),
                         "drop": [float("%.3g" % d) for d in h],
                         "min_drop": float("%.3g" % min(h)),
                         "steps_to_target": (None if first is None
                                             else (first + 1) * chunk),
                         "nonfinite": k in nonfinite})
        out["runs"]["%g" % K] = {
            "variant": {1: "split", 2: "fused", 3: "fused_face"}[variant],
            "seconds": el, "reached_target": sum(
                r["steps_to_target"] is not None for r in reqs),
            "requests": reqs}
        b.close()
    return out


def batched_run(n_inst, cells, steps, hbm_gb, baseline=True):
    """Batched explicit Euler: n_inst seeded jittered Delaunay meshes (cells
    log-uniform in the range), 4-quadrant Riemann states, HLLC + BJ, RK3
    (nRKstages=3, timeOrder=2), CFL 0.9, transmissive boundaries, FP32 flux
    with FP64 state, all in one arena with graph-captured steps and the
    autotuner's operator; then every instance one at a time on the plain
    FP64 single-mesh path."""
    import concurrent.futures as cf
    import multiprocessing as mp
    from ccfd_amd import batch as BT
    from ccfd_amd import gpu as G
    from ccfd_amd import workloads as W
    specs = W.batched_specs(n_inst, seed=4, cells=cells)
    t0 = time.time()
    # spawned workers, not forked ones: this process may already hold a HIP
    # runtime (pinned staging pages that a fork turns copy-on-write)
    with cf.ProcessPoolExecutor(min(16, os.cpu_count() or 4),
                                mp_context=mp.get_context("spawn")) as ex:
        raws = list(ex.map(W.batched_raw, specs, chunksize=2))
    with cf.ThreadPoolExecutor(8) as ex:
        cases = list(ex.map(lambda sr: W.batched_case(*sr), zip(specs, raws)))
    del raws
    ncell = sum(m.n_elems for m, _ in cases)
    log("batched: %d meshes, %d cells, built in %.1f s"
        % (len(cases), ncell, time.time() - t0))
    sch32 = G.params(flux=5, order=2, limiter=1, nrk=3, torder=2, cfl=0.9,
                     precision=G.FP32_FLUX)
    sch64 = G.params(flux=5, order=2, limiter=1, nrk=3, torder=2, cfl=0.9)
    b = BT.Batch(sch32, max_sides=3, hbm_bytes=hbm_gb << 30,
                 max_instances=max(64, n_inst))
    t0 = time.time()
    ids = [b.add(m, p) for m, p in cases]
    log("admitted in %.1f s" % (time.time() - t0))
    variant, ms_var = b.autotune(6)
    b.sync()
    w0 = time.perf_counter()
    b.step(steps)
    b.sync()
    el = time.perf_counter() - w0
    out = {"instances": len(ids), "cells": ncell, "steps": steps,
           "cells_range": list(cells),
           "scheme": "HLLC, Barth-Jespersen, RK3 (3 stages, order 2), CFL "
                     "0.9, FP32 flux / FP64 state",
           "autotune_ms_per_step": {"split": ms_var[1], "fused": ms_var[2],
                                    "fused_face": ms_var[3]},
           "variant": {1: "split", 2: "fused", 3: "fused_face"}[variant],
           "seconds": el,
           "cell_stage_updates_per_s": ncell * 3 * steps / el,
           "solves_per_s": len(ids) / el}
    if baseline:
        import numpy as np
        tot, errs, eerrs = 0.0, [], []
        nan32 = nan64 = 0
        choices = {}
        for k, (m, p) in zip(ids, cases):
            s = G.Solver(m, sch64)
            s.set_prim(p)
            # each mesh on its fastest operator, picked before the timed
            # steps as the arena's is (the state is left as it was)
            v, _, _ = s.autotune(3)
            name = {1: "split", 2: "fused", 3: "fused_face"}.get(v, str(v))
            choices[name] = choices.get(name, 0) + 1
            s.sync()
            w0 = time.perf_counter()
            s.advance(steps, 0.0, 1e300, 1e300, residual=True)
            s.sync()
            tot += time.perf_counter() - w0
            q32, q64 = b.prim(k), s.prim()
            nan32 += int(not np.isfinite(q32).all())
            nan64 += int(not np.isfinite(q64).all())
            errs.append(_density_err(q32, q64))
            eerrs.append(_energy_err(q32, q64))
            s.close()
        out["one_at_a_time_fp64"] = {"seconds": tot,
                                     "operator_choices": choices,
                                     "solves_per_s": len(ids) / tot,
                                     "cell_stage_updates_per_s":
                                         ncell * 3 * steps / tot}
        out["speedup_solves_per_s"] = out["solves_per_s"] / (len(ids) / tot)
        out["density_err_vs_fp64"] = _err_summary(errs)
        bad = [k for k, e in enumerate(errs) if not e[0] <= 1e-6]
        out["density_err_vs_fp64"].update({
            "instances_l2_within_1e-6": len(errs) - len(bad),
            "instances": len(errs),
            # instances over the 1e-6 target: failures of the FP32-flux
            # mode at that tolerance (shocks moved by a fraction of a cell)
            "instances_over_1e-6": [
                {"instance": k, "cells": cases[k][0].n_elems,
                 "density_rel_l2": errs[k][0], "density_rel_linf": errs[k][1],
                 "max_mach": _max_mach(cases[k][1])} for k in bad]})
        out["energy_err_vs_fp64"] = _err_summary(eerrs)
        out["nonfinite_states"] = {"batched_fp32": nan32, "single_fp64": nan64}
    b.close()
    return out


def prep_run(sizes=(2e5, 2e6), kinds=("naca", "bump"), repeats=3):
    """Request-time preparation of one served request, host against device:
    host = ccfd_mesh_build (mesh.Mesh) + ccfd_batch_add (build_soa and
    build_halo on the host, uploads, initial state); device = mesh.RawMesh +
    ccfd_batch_add_raw (pairing, geometry, order and halos on the GPU, same
    initial state). Both paths admit the same instance bit for bit. Best of
    `repeats` after one warm-up admission."""
    import numpy as np
    from ccfd_amd import batch as BT
    from ccfd_amd import gpu as G
    from ccfd_amd import mesh as M
    from ccfd_amd import workloads as W
    sch = G.params(flux=9, order=2, limiter=2, venk_k=5.0, nrk=4, torder=2)
    b = BT.Batch(sch, max_sides=4, local_dt=True, hbm_bytes=16 << 30)
    out = {}
    for kind in kinds:
        for n in sizes:
            spec = {"cells": n, "kind": kind, "mach": 0.7, "aoa": 2.0}
            rec = {}
            for path in ("host", "device"):
                best = None
                for r in range(repeats + 1):
                    t0 = time.perf_counter()
                    mesh, prim, cfl = W.served_case(spec, raw=path == "device")
                    t1 = time.perf_counter()
                    k = (b.add_raw if path == "device" else b.add)(mesh, prim,
                                                                   cfl=cfl)
                    b.sync()
                    t2 = time.perf_counter()
                    b.remove(k)
                    if r and (best is None or t2 - t0 < best[0] + best[1]):
                        best = (t1 - t0, t2 - t1)
                rec[path] = {"prepare_s": best[0], "admit_s": best[1],
                             "total_s": best[0] + best[1]}
            rec["cells"] = int(mesh.n_elems)
            rec["speedup"] = rec["host"]["total_s"] / rec["device"]["total_s"]
            out["%s_%g" % (kind, n)] = rec
            log("prep %s %g: %s" % (kind, n, json.dumps(rec)))
    b.close()
    return out


def arena_ops_run(cells=2e6, steps=40):
    """One served-size request alone in an arena: ms per step (four stages,
    local time steps, AUSMD + VK K=5) of each operator implementation, FP64
    and FP32 flux, on a NACA O-grid and a channel bump. Host clock around
    `steps` steps after 4 warm-up steps, synchronised on both sides."""
    from ccfd_amd import batch as BT
    from ccfd_amd import gpu as G
    from ccfd_amd import workloads as W
    out = {}
    for kind in ("naca", "bump"):
        spec = {"cells": cells, "kind": kind, "mach": 0.7, "aoa": 2.0}
        for prec, pname in ((G.FP64, "fp64"), (G.FP32_FLUX, "fp32_flux")):
            sch = G.params(flux=9, order=2, limiter=2, venk_k=5.0, nrk=4,
                           torder=2, precision=prec)
            b = BT.Batch(sch, max_sides=4, local_dt=True, hbm_bytes=8 << 30)
            mesh, prim, cfl = W.served_case(spec, raw=True)
            b.add_raw(mesh, prim, cfl=cfl)
            res = {}
            for var, vn in ((G.KV_SPLIT, "split"), (G.KV_FUSED, "fused"),
                            (G.KV_FUSED_FACE, "fused_face")):
                b.set_variant(var)
                b.step(4)
                b.sync()
                t0 = time.perf_counter()
                b.step(steps)
                b.sync()
                res[vn] = 1e3 * (time.perf_counter() - t0) / steps
            out["%s_%s" % (kind, pname)] = res
            log("arena ops %s %s: %s" % (kind, pname, json.dumps(res)))
            b.close()
    return out


def implicit_run(steps, cfl, ni=512, nj=256, venk_k=1000.0, chunk=500,
                 modes=("fp64", "fp32_flux"), save_dir=None):
    """Implicit Navier-Stokes around a NACA 2412 (m=0.02, p=0.4, t=0.12) on
    an ni x nj O-grid (outer radius 20 chords, wall spacing 1e-4): Mach 0.5,
    AoA 2 deg, Re 5000 (mu = rho u c / Re), Pr 0.72, Roe flux, 2nd order
    with VK (K=1000: the flow is subsonic and shock-free, and with K=5 the
    limiter's switching stalls the residual near 1e-2), adiabatic no-slip
    wall, characteristic far field. Runs in calls of `chunk` steps (one
    progress line each) up to `steps` or the 1e-8 drop.
    Backward-Euler pseudo-time steps with local steps (CFL driven by the
    GMRES convergence, up to the given value), one Newton iteration per
    step, GMRES (Krylov dimension 30) preconditioned by 4 multicolour
    Gauss-Seidel sweeps on the first-order Jacobian, in FP64 and in
    FP32-flux mode (FP32 Krylov basis, Jacobian and sweeps; FP64 Newton
    residual, Frechet difference, Arnoldi and state)."""
    import math
    import numpy as np
    from ccfd_amd import gpu as G
    from ccfd_amd import mesh as M
    mach, aoa, re = 0.5, 2.0, 5000.0
    t0 = time.time()
    raw = M.naca_ogrid(ni, nj, wall_bc=201, far_bc=501)
    mesh = M.Mesh(raw, [M.BC(201), M.BC(501, rho=1.0, mach=mach, alpha=aoa,
                                         pressure=1.0)])
    c = math.sqrt(1.4)
    a = math.radians(aoa)
    prim = np.tile([1.0, mach * c * math.cos(a), mach * c * math.sin(a), 1.0],
                   (mesh.n_elems, 1))
    mu = mach * c / re
    log("implicit: %d cells, mesh %.1f s" % (mesh.n_elems, time.time() - t0))
    out = {"cells": mesh.n_elems, "max_steps": steps, "cfl_max": cfl,
           "mu": mu, "venk_k": venk_k,
           "scheme": "NS, Roe, 2nd order VK, backward Euler with local "
                     "pseudo-time steps (CFL driven by GMRES convergence), "
                     "1 Newton iteration / step, GMRES nKdim 30 (tolerance "
                     "1e-2), 4 multicolour Gauss-Seidel sweeps on the "
                     "first-order Jacobian"}
    states = {}
    for name, prec in (("fp64", G.FP64), ("fp32_flux", G.FP32_FLUX)):
        if name not in modes:
            continue
        prm = G.params(flux=2, order=2, limiter=2, venk_k=venk_k,
                       navierstokes=1, mu=mu, Pr=0.72, precision=prec)
        s = G.Solver(mesh, prm)
        s.set_prim(prim)
        cfg = G.ImplicitCfg(30, 1, 2, 1e-3, 1e-2, 0.9, 1, 4, 1)
        s.sync()
        w0 = time.perf_counter()
        r0, st, ng, hist = None, 0, 0, []
        while st < steps:
            r, r0, k, _, g = s.implicit_steady(cfg, cfl, min(chunk, steps - st),
                                               1e-8, r0)
            st += k
            ng += g
            hist.append([st, float(r[0] / r0[0])])
            log("implicit %s: %d steps, drop %.3e, %.1f s" % (
                name, st, r[0] / r0[0], time.perf_counter() - w0))
            if r[0] <= 1e-8 * r0[0]:
                break
        s.sync()
        el = time.perf_counter() - w0
        states[name] = s.prim()
        if save_dir:
            np.save(os.path.join(save_dir, "implicit_%s.npy" % name),
                    states[name])
        out[name] = {"seconds": el, "ms_per_step": 1e3 * el / st,
                     "steps": st, "gmres_per_step": ng / st,
                     "residual_drop": float(r[0] / r0[0]),
                     "reached_1e-8": bool(r[0] <= 1e-8 * r0[0]),
                     "history": hist}
        log("implicit %s: %s" % (name, json.dumps(out[name])))
        s.close()
    if len(states) == 2:
        a, b = states["fp32_flux"], states["fp64"]
        for name, f in (("density", _density_err), ("energy", _energy_err)):
            l2, linf = f(a, b)
            out["fp32_vs_fp64_%s_rel_l2" % name] = l2
            out["fp32_vs_fp64_%s_rel_linf" % name] = linf
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=1415,
                    help="points per side of the jittered grid")
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--flux", type=int, default=5)
    ap.add_argument("--venk-k", type=float, default=5.0)
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--variant", choices=["auto", "split", "fused",
                                          "fused_face"],
                    default="auto", help="operator implementation")
    ap.add_argument("--tile", type=int, default=0, choices=[0, 256, 512],
                    help="elements per block of the face-once kernel (0: "
                         "automatic)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also time a 1 GiB streaming copy (reachable HBM "
                         "bandwidth; counter calibration under rocprofv3)")
    ap.add_argument("--workload", choices=["operator", "batched", "served",
                                           "implicit", "prep", "arena-ops",
                                           "served-convergence"],
                    default="operator",
                    help="operator: the 4M-triangle stage benchmark (plus a "
                         "short served sample unless --no-served); batched: "
                         "many meshes in one arena; served: the request "
                         "loop; served-convergence: residual histories of "
                         "the served requests")
    ap.add_argument("--venk-list", default="5,100",
                    help="served-convergence: Venkatakrishnan constants")
    ap.add_argument("--chunk", type=int, default=1000,
                    help="served-convergence: steps between residual reads")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-flux-variants", action="store_true",
                    help="operator workload: only the main flux and precision")
    ap.add_argument("--variant-steps", type=int, default=10,
                    help="timed steps of the other fluxes / FP32 mode")
    ap.add_argument("--served-requests", type=int, default=12)
    ap.add_argument("--served-rate", type=float, default=4.0)
    ap.add_argument("--served-variant", default="auto",
                    choices=["auto", "split", "fused", "fused_face"],
                    help="served workload: operator of both arenas (auto: "
                         "the arena's default)")
    ap.add_argument("--served-venk-k", type=float, default=5.0,
                    help="Venkatakrishnan constant of the served solves")
    ap.add_argument("--requests", type=int, default=2000)
    ap.add_argument("--rate", type=float, default=20.0)
    ap.add_argument("--min-cells", type=float, default=None)
    ap.add_argument("--max-cells", type=float, default=None)
    ap.add_argument("--max-steps", type=int, default=20000)
    ap.add_argument("--instances", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=200)
    ap.add_argument("--hbm-gb", type=int, default=64)
    ap.add_argument("--modes", default="fp64,fp32_flux",
                    help="implicit workload: precisions to run")
    ap.add_argument("--save-state", default=None,
                    help="implicit workload: directory for the final states")
    ap.add_argument("--no-baseline", action="store_true",
                    help="batched workload: skip the one-at-a-time FP64 "
                         "solves (profiling)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: there is no multi-GPU path (a "
              "large mesh would shard as contiguous space-filling-curve "
              "ranges with a 2-ring halo exchange per stage); run with "
              "--gpus 1", file=sys.stderr)
        sys.exit(2)

    from ccfd_amd import gpu as G

    if args.workload == "batched":
        cells = (args.min_cells or 2e3, args.max_cells or 5e5)
        if args.copy_probe:
            G.copy_probe(1 << 30, 20)
        r = batched_run(args.instances, cells, args.batch_steps, args.hbm_gb,
                        baseline=not args.no_baseline)
        out = {"metric": "batched cell-stage updates/s (many meshes, one "
                         "launch per stage)",
               "value": r["cell_stage_updates_per_s"],
               "unit": "cell-stage updates/s", "n_gpus": 1,
               "steps": args.batch_steps, "warmup": 0,
               "higher_is_better": True, "scaling": "weak",
               "vs_baseline": None, "dtype": "f32 flux / f64 state",
               "data": "synthetic (seeded Delaunay meshes, 4-quadrant "
                       "Riemann states)",
               "config": {"workload": "%d instances, %d cells, %d steps"
                          % (r["instances"], r["cells"], r["steps"])},
               "batched": r}
        print(json.dumps(out), flush=True)
        return
    if args.workload == "implicit":
        modes = args.modes.split(",")
        r = implicit_run(args.batch_steps if args.batch_steps != 200 else 10000,
                         1e4, modes=modes, save_dir=args.save_state)
        main_mode = "fp32_flux" if "fp32_flux" in modes else modes[0]
        out = {"metric": "implicit pseudo-time steps/s (NACA O-grid NS, "
                         "JFNK-GMRES + LU-SGS)",
               "value": 1e3 / r[main_mode]["ms_per_step"],
               "unit": "steps/s", "n_gpus": 1,
               "steps": r[main_mode]["steps"],
               "warmup": 0, "higher_is_better": True, "scaling": "weak",
               "vs_baseline": None,
               "dtype": ("f32 flux / f64 state" if main_mode == "fp32_flux"
                         else "f64"),
               "data": "synthetic (analytic NACA 2412 O-grid, free stream)",
               "config": {"workload": "512x256 O-grid NS Re 5000, to a 1e-8 "
                                      "drop or %d steps" % r["max_steps"]},
               "implicit": r}
        print(json.dumps(out), flush=True)
        return
    if args.workload == "arena-ops":
        r = arena_ops_run()
        out = {"metric": "ms per step of one 2M-cell request in an arena",
               "value": r["naca_fp32_flux"]["split"], "unit": "ms",
               "n_gpus": 1, "steps": 40, "warmup": 4,
               "higher_is_better": False, "scaling": "weak",
               "vs_baseline": None, "dtype": "f64 / f32 flux",
               "data": "synthetic (NACA O-grid and channel-bump requests)",
               "config": {"workload": "one 2e6-cell served request, each "
                                      "operator implementation"},
               "arena_ops": r}
        print(json.dumps(out), flush=True)
        return
    if args.workload == "prep":
        r = prep_run()
        out = {"metric": "request preparation s (host vs device)",
               "value": r["naca_2e+06"]["device"]["total_s"], "unit": "s",
               "n_gpus": 1, "steps": None, "warmup": 1,
               "higher_is_better": False, "scaling": "weak",
               "vs_baseline": None, "dtype": "f64",
               "data": "synthetic (NACA O-grid and channel-bump requests)",
               "config": {"workload": "one served request's mesh set-up and "
                                      "admission, 2e5 and 2e6 cells"},
               "prep": r}
        print(json.dumps(out), flush=True)
        return
    if args.workload == "served-convergence":
        cells = (args.min_cells or 1e4, args.max_cells or 2e6)
        ks = [float(k) for k in args.venk_list.split(",")]
        r = served_convergence_run(args.served_requests, cells, ks,
                                   max(1, args.max_steps // args.chunk),
                                   args.chunk, args.hbm_gb)
        first = r["runs"]["%g" % ks[0]]
        out = {"metric": "served requests reaching a 1e-6 density residual "
                         "drop (explicit local steps, one FP64 arena)",
               "value": first["reached_target"], "unit": "requests",
               "n_gpus": 1, "steps": r["steps"], "warmup": 0,
               "higher_is_better": True, "scaling": "weak",
               "vs_baseline": None, "dtype": "f64",
               "data": "synthetic (channel-with-bump and NACA O-grid "
                       "requests)",
               "config": {"workload": "%d requests, cells %g..%g, %d steps"
                          % (args.served_requests, cells[0], cells[1],
                             r["steps"])},
               "served_convergence": r}
        print(json.dumps(out), flush=True)
        return
    if args.workload == "served":
        cells = (args.min_cells or 1e4, args.max_cells or 2e6)
        from ccfd_amd import gpu as G
        var = {"auto": None, "split": G.KV_SPLIT, "fused": G.KV_FUSED,
               "fused_face": G.KV_FUSED_FACE}[args.served_variant]
        r = served_run(args.requests, cells, args.rate, args.max_steps,
                       args.hbm_gb, venk_k=args.served_venk_k,
                       baseline=not args.no_baseline, variant=var)
        out = {"metric": "served steady solves/s (p50/p99 latency)",
               "value": r["served"]["solves_per_s"], "unit": "solves/s",
               "n_gpus": 1, "steps": None, "warmup": 0,
               "higher_is_better": True, "scaling": "weak",
               "vs_baseline": None, "dtype": "f32 flux / f64 state",
               "data": "synthetic (channel-with-bump and NACA O-grid "
                       "requests, Poisson arrivals)",
               "config": {"workload": "%d requests, cells %g..%g, rate %g/s"
                          % (args.requests, cells[0], cells[1], args.rate)},
               "served": r}
        print(json.dumps(out), flush=True)
        return

    raw, bcs, mesh, prim = build_case(args.n, args.seed)
    prec = G.FP64 if args.precision == "fp64" else G.FP32_FLUX

    def prm_of(flux, precision):
        return G.params(flux=flux, order=2, limiter=2, venk_k=args.venk_k,
                        nrk=4, torder=2, cfl=0.9, precision=precision)
    t0 = time.time()
    solver = G.Solver(mesh, prm_of(args.flux, prec))
    solver.set_variant({"auto": G.KV_AUTO, "split": G.KV_SPLIT,
                        "fused": G.KV_FUSED,
                        "fused_face": G.KV_FUSED_FACE}[args.variant])
    solver.set_tile(args.tile)
    tuned = None
    if args.variant == "auto" and args.tile == 0:
        # the autotuner times the operator implementations and tiles on the
        # workload's own state and keeps the fastest
        solver.set_prim(prim)
        _, _, tuned = solver.autotune(10)
        log("autotune (ms per operator call): %s" % json.dumps(tuned))
    kv = solver.variant()
    tile = solver.tile()
    fused = kv in (G.KV_FUSED, G.KV_FUSED_FACE)
    solver.set_prim(prim)
    log("upload %.1f s" % (time.time() - t0))
    N = mesh.n_elems
    big = 1e300
    uts = {}   # (flux, precision) -> dU/dt of the initial state

    def measure(flux, precision, steps, warmup):
        """dU/dt of the initial state, then `warmup` untimed and `steps`
        timed steps (calcTimeStep + 4 stages + globalResidual each)"""
        solver.set_params(prm_of(flux, precision))
        solver.set_prim(prim)
        solver.fvtd(0.0)
        uts[(flux, precision)] = solver.ut()
        solver.set_prim(prim)
        if warmup > 0:
            solver.advance(warmup, 0.0, big, big, residual=True)
        solver.sync()
        w0 = time.perf_counter()
        ms, launches, res = solver.profile_steps(steps, big, big,
                                                 residual=True)
        solver.sync()
        el = time.perf_counter() - w0
        bstage = B_STAGE_FP64 if precision == G.FP64 else B_STAGE_FP32
        bker = bstage if fused else (B_FLUX_FP64 if precision == G.FP64
                                     else B_FLUX_FP32)
        achieved = bker * N / (ms["flux"] * 1e-3) / 1e9
        return {"value": N * 4 * steps / el, "ms_per_step": 1e3 * el / steps,
                "seconds": el, "kernel_ms": {k: round(v, 4)
                                             for k, v in ms.items()},
                "launches": launches, "residual": res.tolist(),
                "achieved_gbs": achieved, "bytes_per_cell": bker,
                "frac": achieved / PEAK_HBM_GBS}

    # timed region of the headline: K steps, HIP events around every kernel
    main = measure(args.flux, prec, args.steps, args.warmup)
    ms = main["kernel_ms"]
    log("main: %.3f ms / step, kernel %.4f ms" % (main["ms_per_step"],
                                                   ms["flux"]))
    pname = args.precision
    if fused:
        kname = ("%s (reconstruction + limiter + Riemann flux + gather + "
                 "RK stage update, side states in LDS)"
                 % ("fused_face_kernel" if kv == G.KV_FUSED_FACE
                    else "fused_kernel"))
        kkey = "%s_%s" % ("fused_face" if kv == G.KV_FUSED_FACE else "fused",
                          pname)
    else:
        kname = "flux_kernel (Riemann flux + gather + RK stage update)"
        kkey = "flux_%s" % pname
    traffic, tsrc = pmc_traffic(kkey)
    roof = {"bound": "hbm", "achieved": round(main["achieved_gbs"], 1),
            "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": round(main["frac"], 4),
            "traffic": traffic,
            "traffic_source": tsrc,
            "kernel": kname,
            "kernel_ms": ms["flux"],
            "kernel_launches": main["launches"][2],
            "bytes_per_cell": main["bytes_per_cell"],
            "bytes_note": "achieved = algorithmic bytes per cell and stage "
                          "(the two-phase dataflow of fvTimeDerivative + RK "
                          "update: %g B) x cells / mean kernel time over the "
                          "timed launches" % main["bytes_per_cell"]}
    if traffic:
        roof["measured_bytes_per_cell"] = round(traffic / N, 1)
        roof["measured_hbm_gbs"] = round(traffic / (ms["flux"] * 1e-3) / 1e9, 1)
    if fused:
        bmin = fused_min_bytes(prec == G.FP64, 4)
        roof["compulsory_bytes_per_cell"] = bmin
        roof["compulsory_gbs"] = round(bmin * N / (ms["flux"] * 1e-3) / 1e9, 1)

    # the other fluxes of the workload, and the FP32-flux mode
    others = {}
    flux_list = [] if args.no_flux_variants else [
        f for f in (2, 3, 5, 9) if f != args.flux]
    for f in flux_list:
        r = measure(f, prec, args.variant_steps, 2)
        log("%s: %.3f ms / step" % (FLUX_NAMES.get(f, f), r["ms_per_step"]))
        others[FLUX_NAMES.get(f, str(f))] = r
    fp32 = None
    if prec == G.FP64 and not args.no_flux_variants:
        fp32 = measure(args.flux, G.FP32_FLUX, args.variant_steps, 2)
        log("fp32 flux: %.3f ms / step" % fp32["ms_per_step"])

    copy_gbs = None
    if args.copy_probe:
        copy_gbs = G.copy_probe(1 << 30, 20)
        log("copy probe: %.0f GB/s" % copy_gbs)
    solver.close()
    del solver

    # the reference's fvTimeDerivative on the same mesh and state: parity of
    # every flux, and the CPU baseline
    cpu, refut = None, {}
    if not args.no_cpu:
        try:
            cpu, refut = reference_in_child(
                raw, bcs, prim, args.flux, args.venk_k, args.cpu_seconds,
                [args.flux] + flux_list)
        except Exception as e:   # must not hide the GPU result
            log("reference run failed:", e)
    del raw, bcs
    parity = {}
    for (f, p), ut in uts.items():
        if f not in refut:
            continue
        key = "%s_%s" % (FLUX_NAMES.get(f, str(f)),
                         "fp64" if p == G.FP64 else "fp32_flux")
        parity[key] = _ut_parity(ut, refut[f])
    del uts, refut
    for name, r in others.items():
        r["parity_vs_reference"] = parity.get("%s_%s" % (name, pname))

    out = {
        "metric": "cell-stage updates/s (4M tri HLLC, 2nd order VK, RK4)",
        "value": main["value"],
        "unit": "cell-stage updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == G.FP64 else "f32 flux / f64 state",
        "data": "synthetic (seeded jittered Delaunay mesh, seeded smooth state)",
        "config": {
            "workload": "fvTimeDerivative + RK stage, periodic jittered "
                        "Delaunay %d triangles, HLLC, VK K=%g, "
                        "nRKstages=4 timeOrder=2, %s; a step is "
                        "calcTimeStep + 4 stages + globalResidual"
                        % (N, args.venk_k, args.precision),
            "cells": N, "stages_per_step": 4, "flux": args.flux,
            "precision": args.precision,
            "operator": {G.KV_SPLIT: "split", G.KV_FUSED: "fused",
                         G.KV_FUSED_FACE: "fused_face"}[kv],
            "tile_elements": tile,
            "autotune_ms_per_call": tuned,
        },
        "kernel_ms": ms,
        "residual_last_step": main["residual"],
        "parity_vs_reference": parity.get("%s_%s" % (
            FLUX_NAMES.get(args.flux, str(args.flux)), pname)),
        "fluxes": others,
        "fp32_flux": fp32,
        "copy_probe_gbs": copy_gbs,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if fp32 is not None:
        fp32["roofline_frac_330B"] = fp32.pop("frac")
        fp32["error_vs_reference_fp64"] = parity.get(
            "%s_fp32_flux" % FLUX_NAMES.get(args.flux, str(args.flux)))
    if not args.no_batched:
        try:
            out["batched_sample"] = batched_run(
                args.instances, (args.min_cells or 2e3,
                                 args.max_cells or 5e5), args.batch_steps,
                args.hbm_gb)
        except Exception as e:   # must not hide the main result
            log("batched sample failed:", e)
            out["batched_sample"] = {"error": str(e)}
    if not args.no_served:
        # a short served sample over the full request size range
        try:
            out["served_sample"] = served_run(
                args.served_requests, (args.min_cells or 1e4,
                                       args.max_cells or 2e6),
                args.served_rate, min(args.max_steps, 8000), 32,
                venk_k=args.served_venk_k)
            out["served_sample"]["note"] = (
                "%d requests (the full stream has 2000) over the full 1e4 - "
                "2e6 cell range" % args.served_requests)
        except Exception as e:   # must not hide the main result
            log("served sample failed:", e)
            out["served_sample"] = {"error": str(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
