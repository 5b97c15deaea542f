#!/usr/bin/env python3
"""Benchmark: explicit RK stages of ccfd's finite-volume operator on one
MI355X.

Workload (default): Delaunay triangulation of a jittered 1415 x 1415 point
grid on the periodic unit square (3,998,792 triangles), smooth seeded state
(density / pressure = 1 + 0.1 x 4 sine modes, Mach 0.5 flow at a random
angle), HLLC flux, 2nd order with Venkatakrishnan limiting (K = 5),
nRKstages = 4 / timeOrder = 2, FP64. A step is one calcTimeStep plus four
fvTimeDerivative + RK stage updates; the metric counts cell-stage updates.

Prints one JSON line (rank 0) with the metric, a roofline entry for the
dominant kernel (the fused operator + stage update, or the flux + gather +
stage update kernel of the split pair) from HIP-event kernel timings over the
timed region, and a CPU baseline: the reference's own OpenMP
fvTimeDerivative on the same mesh and state, timed on this host.
"""
import argparse
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
# what the fused kernel itself must move per element and stage (each array
# once, side states in LDS): state 32 + links 12 + LSQ weights 48 + GP 48 +
# eps^2 8 + side normals 48 + signed lengths 24 + 1/area 8 + U0 32 + U 32 +
# primitive out 32 = 324 B (+32 B dU/dt on the last stage) in FP64; halo
# re-reads come on top
B_FUSED_MIN_FP64 = 324.0
B_FUSED_MIN_FP32 = 200.0


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


def cpu_baseline(mesh, prim, flux, venk_k, seconds):
    """The reference's fvTimeDerivative (compiled from its own sources with
    OpenMP, oracle/_ref/libccfdref_omp.so) on the same mesh and state."""
    from oracle import bindings as B
    if not B.ref_available(False) or not os.path.exists(
            os.path.join(ROOT, "oracle", "_ref", "libccfdref_omp.so")):
        return None
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(
        16, os.cpu_count() or 1)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    ini = "/tmp/ccfd_bench_ref_%d.ini" % os.getpid()
    B.write_ini(ini, [("fluxFunction", flux), ("nBC", 4),
                      ("BCtype", 701), ("connection", (0.0, 1.0)),
                      ("BCtype", 702), ("connection", (0.0, 0.0)),
                      ("BCtype", 711), ("connection", (1.0, 0.0)),
                      ("BCtype", 712), ("connection", (0.0, 0.0)),
                      ("spatialOrder", 2), ("limiter", 2),
                      ("venk_k", venk_k)])
    devnull = os.open(os.devnull, os.O_WRONLY)
    saved = os.dup(1)
    os.dup2(devnull, 1)   # the reference prints its parameter echo
    try:
        ref = B.Ref(omp=True)
        ref.setup(ini, mesh)
        ref.set_prim(prim)
        ref.fvtd_only(0.0)   # warm-up
        calls, t0 = 0, time.time()
        while calls < 3 or time.time() - t0 < seconds:
            ref.fvtd_only(0.0)
            calls += 1
        el = time.time() - t0
    finally:
        os.dup2(saved, 1)
        os.close(devnull)
        os.close(saved)
    try:
        os.unlink(ini)
    except OSError:
        pass
    cpu = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {
        "value": mesh.n_elems * calls / el,
        "unit": "cell-evaluations/s (reference fvTimeDerivative, OpenMP)",
        "cores": threads,
        "kind": "reference",
        "sample": "%d fvTimeDerivative calls on the same %d-triangle mesh "
                  "and state (%.1f s; RK update ~2%% of CPU time not "
                  "included), CPU: %s" % (calls, mesh.n_elems, el, cpu),
    }


def pmc_traffic(kernel_key):
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 --pmc summary (FETCH_SIZE / WRITE_SIZE, copy-kernel scaled)."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
        return d.get(kernel_key)
    except (OSError, ValueError):
        return None


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
    ap.add_argument("--variant", choices=["auto", "split", "fused"],
                    default="auto", help="operator implementation")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also time a 1 GiB streaming copy (reachable HBM "
                         "bandwidth; counter calibration under rocprofv3)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: there is no multi-GPU path (a "
              "large mesh would shard as contiguous space-filling-curve "
              "ranges with a 2-ring halo exchange per stage); run with "
              "--gpus 1", file=sys.stderr)
        sys.exit(2)

    from ccfd_amd import gpu as G

    raw, bcs, mesh, prim = build_case(args.n, args.seed)
    prec = G.FP64 if args.precision == "fp64" else G.FP32_FLUX
    prm = G.params(flux=args.flux, order=2, limiter=2, venk_k=args.venk_k,
                   nrk=4, torder=2, cfl=0.9, precision=prec)
    t0 = time.time()
    solver = G.Solver(mesh, prm)
    solver.set_variant({"auto": G.KV_AUTO, "split": G.KV_SPLIT,
                        "fused": G.KV_FUSED}[args.variant])
    fused = solver.variant() == G.KV_FUSED
    solver.set_prim(prim)
    log("upload %.1f s" % (time.time() - t0))
    big = 1e300
    t = 0.0
    if args.warmup > 0:
        t, _ = solver.advance(args.warmup, t, big, big, residual=False)
    solver.sync()

    # timed region: K steps, HIP events around every kernel on the stream
    solver.sync()
    w0 = time.perf_counter()
    ms, launches = solver.profile_steps(args.steps, big, big)
    solver.sync()
    w1 = time.perf_counter()
    el = w1 - w0
    stages = 4 * args.steps
    N = mesh.n_elems
    value = N * stages / el
    ms_step = 1e3 * el / args.steps

    bstage = B_STAGE_FP64 if prec == G.FP64 else B_STAGE_FP32
    if fused:
        bker = bstage
        kname = ("fused_kernel (reconstruction + limiter + Riemann flux + "
                 "gather + RK stage update, side states in LDS)")
        kkey = "fused_%s" % args.precision
    else:
        bker = B_FLUX_FP64 if prec == G.FP64 else B_FLUX_FP32
        kname = "flux_kernel (Riemann flux + gather + RK stage update)"
        kkey = "flux_%s" % args.precision
    achieved = bker * N / (ms["flux"] * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 1),
            "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBS, 4),
            "traffic": pmc_traffic(kkey),
            "kernel": kname,
            "kernel_ms": round(ms["flux"], 4),
            "bytes_per_cell": bker}
    if fused:
        bmin = B_FUSED_MIN_FP64 if prec == G.FP64 else B_FUSED_MIN_FP32
        roof["compulsory_bytes_per_cell"] = bmin
        roof["compulsory_gbs"] = round(bmin * N / (ms["flux"] * 1e-3) / 1e9, 1)

    copy_gbs = None
    if args.copy_probe:
        copy_gbs = G.copy_probe(1 << 30, 20)
        log("copy probe: %.0f GB/s" % copy_gbs)

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(mesh, prim, args.flux, args.venk_k,
                               args.cpu_seconds)
        except Exception as e:   # the baseline must not hide the GPU result
            log("cpu baseline failed:", e)
            cpu = None

    out = {
        "metric": "cell-stage updates/s (4M tri HLLC, 2nd order VK, RK4)",
        "value": value,
        "unit": "cell-stage updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == G.FP64 else "f32 flux / f64 state",
        "data": "synthetic (seeded jittered Delaunay mesh, seeded smooth state)",
        "config": {
            "workload": "fvTimeDerivative + RK stage, periodic jittered "
                        "Delaunay %d triangles, HLLC, VK K=%g, "
                        "nRKstages=4 timeOrder=2, %s"
                        % (N, args.venk_k, args.precision),
            "cells": N, "stages_per_step": 4, "flux": args.flux,
            "precision": args.precision,
            "operator": "fused" if fused else "split",
        },
        "hbm_gbs_stage_effective": round(value * bstage / 1e9, 1),
        "kernel_ms": {k: round(v, 4) for k, v in ms.items()},
        "copy_probe_gbs": copy_gbs,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    solver.close()


if __name__ == "__main__":
    main()
