"""Synthetic workloads of the batched and served paths.

batched_specs  many independent explicit Euler runs: seeded jittered
               Delaunay meshes on the unit square with cell counts
               log-uniform in a range, transmissive boundaries (401), and a
               4-quadrant 2-D Riemann problem as initial state.
served_specs   steady solves: a channel with a bump or a NACA 4-digit O-grid
               (probability 1/2 each), cell count log-uniform in a range,
               free-stream Mach ~ U[0.3, 2.5] and angle of attack
               ~ U[-4, 8] deg (channel: along the axis), characteristic
               far field / inflow (501), slip walls (101), pressure outlet
               (801); initial state = free stream.
"""
import math

import numpy as np

from . import batch as BT
from . import mesh as M


def _logu(rng, lo, hi, n=None):
    return np.exp(rng.uniform(math.log(lo), math.log(hi), n))


# ------------------------------------------------------------- batched runs

def batched_specs(n_inst=256, seed=4, cells=(2e3, 5e5)):
    """(points per side, mesh seed, state seed, target cells) per instance"""
    rng = np.random.Generator(np.random.PCG64(seed))
    target = np.rint(_logu(rng, cells[0], cells[1], n_inst)).astype(int)
    specs = []
    for k, n_cells in enumerate(target):
        pts = int(round(math.sqrt(n_cells / 2.0))) + 1   # 2 (p - 1)^2 cells
        specs.append({"k": k, "points": max(pts, 4), "mesh_seed": seed * 1000 + k,
                      "state_seed": seed * 7919 + k, "cells": int(n_cells)})
    return specs


def batched_raw(spec):
    """mesh connectivity of one batched instance (picklable; heavy part)"""
    raw, bcs = M.jittered_delaunay(spec["points"], 0.3, spec["mesh_seed"],
                                   periodic=False, bc_code=401)
    return raw, [b.code for b in bcs]


def batched_case(spec, raw=None):
    """(mesh, initial primitive state) of one batched instance"""
    if raw is None:
        raw, codes = batched_raw(spec)
    else:
        raw, codes = raw
    mesh = M.Mesh(raw, [M.BC(c) for c in codes])
    return mesh, BT.riemann4_state(mesh.bary, spec["state_seed"])


# --------------------------------------------------------------- served

def served_specs(n_req=2000, seed=5, cells=(1e4, 2e6)):
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for j in range(n_req):
        kind = "naca" if rng.uniform() < 0.5 else "channel"
        n = float(_logu(rng, cells[0], cells[1]))
        mach = float(rng.uniform(0.3, 2.5))
        aoa = float(rng.uniform(-4.0, 8.0))
        out.append({"id": j, "kind": kind, "cells": int(round(n)),
                    "mach": mach, "aoa": aoa if kind == "naca" else 0.0})
    return out


def served_case(spec, gamma=1.4, raw=False):
    """(mesh, free-stream primitive state, CFL) of one steady solve; with raw
    the mesh is a mesh.RawMesh (preprocessed on the device at admission)
    instead of a host-preprocessed mesh.Mesh"""
    n, mach, aoa = spec["cells"], spec["mach"], spec["aoa"]
    if spec["kind"] == "naca":
        nj = max(8, int(round(math.sqrt(n / 2.0))))
        ni = max(16, 2 * nj)
        raw_mesh = M.naca_ogrid(ni, nj, wall_bc=101, far_bc=501)
        bcs = [M.BC(101), M.BC(501, rho=1.0, mach=mach, alpha=aoa,
                                pressure=1.0, gamma=gamma)]
    else:
        nj = max(4, int(round(math.sqrt(n / 3.0))))
        ni = 3 * nj
        raw_mesh = M.channel_bump(ni, nj, wall_bc=101, in_bc=501,
                                  out_bc=801)
        bcs = [M.BC(101), M.BC(501, rho=1.0, mach=mach, alpha=0.0,
                                pressure=1.0, gamma=gamma),
               M.BC(801, pressure=1.0)]
    mesh = M.RawMesh(raw_mesh, bcs) if raw else M.Mesh(raw_mesh, bcs)
    c = math.sqrt(gamma)
    a = math.radians(aoa)
    prim = np.tile([1.0, mach * c * math.cos(a), mach * c * math.sin(a), 1.0],
                   (mesh.n_elems, 1))
    # local steps at CFL 2 (the 4-stage scheme converges steady solves
    # several times faster there than at 0.9; 3 stalls)
    return mesh, prim, 2.0
