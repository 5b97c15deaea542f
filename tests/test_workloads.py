"""The synthetic workloads of the batched and served paths
(ccfd_amd/workloads.py): seeded and reproducible, with the distributions
the batched (many Delaunay meshes, 4-quadrant Riemann states) and served
(channel / NACA steady solves) workloads are defined by. Host only."""
import math

import numpy as np

from ccfd_amd import workloads as W


def test_batched_specs_are_seeded_and_log_uniform():
    a = W.batched_specs(256, 4, (2e3, 5e5))
    b = W.batched_specs(256, 4, (2e3, 5e5))
    assert a == b
    assert W.batched_specs(256, 5, (2e3, 5e5)) != a
    n = np.array([s["cells"] for s in a])
    assert n.min() >= 2e3 and n.max() <= 5e5
    # log-uniform: the log of the size is uniform over [ln 2e3, ln 5e5]
    u = (np.log(n) - math.log(2e3)) / (math.log(5e5) - math.log(2e3))
    assert abs(u.mean() - 0.5) < 0.06 and abs(u.std() - math.sqrt(1 / 12)) < 0.04
    # points per side give about the target count (2 (p - 1)^2 triangles)
    for s in a[:20]:
        assert abs(2 * (s["points"] - 1) ** 2 - s["cells"]) <= 4 * s["points"]
    assert len({s["mesh_seed"] for s in a}) == 256


def test_batched_case_is_a_four_quadrant_riemann_state():
    spec = W.batched_specs(4, 4, (2e3, 3e3))[0]
    mesh, prim = W.batched_case(spec)
    assert prim.shape == (mesh.n_elems, 4)
    x, y = mesh.bary[:, 0], mesh.bary[:, 1]
    quad = (x > 0.5).astype(int) + 2 * (y > 0.5).astype(int)
    for q in range(4):
        rows = prim[quad == q]
        assert len(rows) > 0
        # one constant state per quadrant, inside the stated ranges
        assert np.all(rows == rows[0])
        assert 0.1 <= rows[0][0] <= 2 and 0.1 <= rows[0][3] <= 2
        assert -1 <= rows[0][1] <= 1 and -1 <= rows[0][2] <= 1
    assert len({tuple(prim[quad == q][0]) for q in range(4)}) == 4
    mesh2, prim2 = W.batched_case(spec)
    assert np.array_equal(prim, prim2) and mesh2.n_elems == mesh.n_elems


def test_served_specs_follow_the_stream_definition():
    s = W.served_specs(2000, 5, (1e4, 2e6))
    assert s == W.served_specs(2000, 5, (1e4, 2e6))
    kinds = np.array([r["kind"] for r in s])
    assert 0.45 < np.mean(kinds == "naca") < 0.55
    mach = np.array([r["mach"] for r in s])
    aoa = np.array([r["aoa"] for r in s])
    n = np.array([r["cells"] for r in s])
    assert mach.min() >= 0.3 and mach.max() <= 2.5
    assert aoa[kinds == "naca"].min() >= -4 and aoa[kinds == "naca"].max() <= 8
    assert np.all(aoa[kinds == "channel"] == 0.0)
    assert n.min() >= 1e4 and n.max() <= 2e6
    # E[N] of a log-uniform [1e4, 2e6] is about 376k
    assert 3.0e5 < n.mean() < 4.5e5


def test_served_case_sizes_and_free_stream():
    for kind in ("naca", "channel"):
        spec = {"id": 0, "kind": kind, "cells": 3000, "mach": 0.8,
                "aoa": 2.0 if kind == "naca" else 0.0}
        mesh, prim, cfl = W.served_case(spec)
        assert 0.5 * 3000 < mesh.n_elems < 2.0 * 3000
        assert cfl > 0
        c = math.sqrt(1.4 * prim[0, 3] / prim[0, 0])
        assert math.isclose(math.hypot(prim[0, 1], prim[0, 2]) / c, 0.8)
        assert np.all(prim == prim[0])
