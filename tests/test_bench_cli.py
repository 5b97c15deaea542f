"""bench.py's command-line contract that needs no GPU: one GPU only (a
request for more exits with a message before anything touches a device),
and the byte accounting the roofline figure is quoted against."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "bench.py")


def _run(args, env=None):
    e = dict(os.environ)
    e.update(env or {})
    return subprocess.run([sys.executable, BENCH] + args, capture_output=True,
                          text=True, timeout=120, env=e, cwd=ROOT)


@pytest.mark.parametrize("n", [2, 8])
def test_more_than_one_gpu_is_refused(n):
    r = _run(["--gpus", str(n)])
    assert r.returncode == 2
    assert "one GPU" in r.stderr and "--gpus 1" in r.stderr
    assert r.stdout == ""


def test_a_distributed_launch_is_refused():
    r = _run([], env={"WORLD_SIZE": "4", "RANK": "0"})
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_help_lists_the_workloads():
    r = _run(["--help"])
    assert r.returncode == 0
    for w in ("operator", "batched", "served", "implicit"):
        assert w in r.stdout


def test_compulsory_bytes_of_the_fused_stage():
    sys.path.insert(0, ROOT)
    import bench
    # FP64: 292 B per element and stage, plus U and dU/dt (64 B) written by
    # the last stage only; FP32 flux: 168 + 48
    assert bench.fused_min_bytes(True, 4) == 308.0
    assert bench.fused_min_bytes(True, 1) == 356.0
    assert bench.fused_min_bytes(False, 4) == 180.0
    assert bench.fused_min_bytes(False, 3) == 184.0


def test_error_summaries_against_fp64():
    sys.path.insert(0, ROOT)
    import bench
    import numpy as np
    b = np.array([[1.0, 0.0, 0.0, 1.0], [2.0, 1.0, 0.0, 0.4]])
    # E = p / 0.4 + rho |v|^2 / 2: 2.5 and 2.0
    a = np.array([[1.0, 0.0, 0.0, 1.1], [2.0, 1.0, 1.0, 0.4]])
    # E of a: 2.75 and 3.0; density unchanged
    l2, linf = bench._energy_err(a, b)
    assert np.isclose(l2, np.hypot(0.25, 1.0) / np.hypot(2.5, 2.0))
    assert np.isclose(linf, 1.0 / 2.5)
    assert bench._density_err(a, b) == (0.0, 0.0)
    assert bench._energy_err(b, b) == (0.0, 0.0)
    errs = [(1e-7, 1e-5), (3e-7, 2e-6), (2e-6, 4e-4)]
    s = bench._err_summary(errs, [True, True, False])
    assert s["l2_max"] == 2e-6 and s["l2_median"] == 3e-7
    assert s["linf_max"] == 4e-4
    assert s["both_converged"] == 2
    assert s["both_converged_l2_max"] == 3e-7
    assert s["both_converged_linf_max"] == 1e-5
    assert "both_converged" not in bench._err_summary(errs)


def test_served_sample_with_a_stand_in_batch(monkeypatch):
    """the served workload's bookkeeping on the host: requests through the
    real serving loop and workload generator, with a stand-in batch whose
    FP32-flux instances return states 1e-7 off; the summary carries both
    throughputs, the latency percentiles and the density and energy errors
    of the FP32 results against FP64"""
    sys.path.insert(0, ROOT)
    import numpy as np
    import bench
    from ccfd_amd import batch as BT
    from ccfd_amd import gpu as G

    class StandIn:
        def __init__(self, scheme, max_sides, local_dt, hbm_bytes):
            self.fp32 = scheme["precision"] == G.FP32_FLUX
            self.cap = hbm_bytes // 2000
            self.inst, self.next = {}, 0

        def fits(self, mesh):
            return sum(len(p) for p, _ in self.inst.values()) + \
                mesh.n_elems <= self.cap

        def add(self, mesh, prim, cfl=0.9):
            k, self.next = self.next, self.next + 1
            self.inst[k] = [np.array(prim, dtype=float), 0]
            return k

        def remove(self, k):
            del self.inst[k]

        def step(self, n):
            for i in self.inst.values():
                i[1] += n

        def poll(self):
            return {k: {"steps": s, "res0": np.ones(4), "nonfinite": False,
                        "res": np.full(4, np.exp(-s / 60.0))}
                    for k, (p, s) in self.inst.items()}

        def prim(self, k):
            p = self.inst[k][0]
            return p * (1.0 + 1e-7) if self.fp32 else p.copy()

        def cons(self, k):
            p = self.prim(k)
            c = p.copy()
            c[:, 1:3] *= p[:, :1]
            c[:, 3] = p[:, 3] / 0.4 + 0.5 * p[:, 0] * (p[:, 1] ** 2 +
                                                       p[:, 2] ** 2)
            return c

        def close(self):
            pass

    monkeypatch.setattr(BT, "Batch", StandIn)
    monkeypatch.setattr(G, "params", lambda **kw: dict(kw))
    out = bench.served_run(6, (1500, 2500), 200.0, 3000, 1)
    assert out["served"]["solves"] == 6
    assert out["one_at_a_time_fp64"]["solves"] == 6
    assert out["served"]["latency_p99_s"] >= out["served"]["latency_p50_s"]
    assert out["speedup_solves_per_s"] > 0
    for key in ("density_err_vs_fp64", "energy_err_vs_fp64"):
        e = out[key]
        assert 0 < e["l2_max"] < 1e-6 and 0 < e["linf_max"] < 1e-6, (key, e)
        assert e["both_converged"] == 6


def test_traffic_figures_come_from_the_committed_profiles():
    """bench.py's roofline.traffic reads profiles/pmc_traffic.json; each
    entry must be the scaled FETCH_SIZE + WRITE_SIZE bytes per dispatch that
    the committed rocprofv3 summary lists for that kernel"""
    import json
    import re
    prof = os.path.join(ROOT, "profiles")
    with open(os.path.join(prof, "pmc_traffic.json")) as f:
        traffic = json.load(f)
    rows = {
        "fused_face_fp64": ("fused_face_kernel_fp64_4M_delaunay.md",
                            "fused_face_kernel<5, double, 2, 3>"),
        "fused_face_fp32": ("fused_face_kernel_fp32_4M_delaunay.md",
                            "fused_face_kernel<5, float, 2, 3>"),
        "batched_fused_fp32": ("fused_kernel_fp32_batched_64_meshes.md",
                               "ccfd::fused_kernel<5, float, 1, 3>"),
    }
    assert set(traffic) == set(rows)
    for key, (md, kernel) in rows.items():
        text = open(os.path.join(prof, md)).read()
        head = text.index("## HBM traffic")
        line = [x for x in text[head:].splitlines() if kernel in x][0]
        cells = [c.strip() for c in line.strip("|").split("|")]
        fetch, write, scaled = float(cells[1]), float(cells[2]), float(cells[3])
        scale = float(re.search(r"FETCH is scaled by ([0-9.]+)",
                                text).group(1))
        assert abs(traffic[key] / scaled - 1) < 1e-3, key
        assert abs((fetch * scale + write) * 1024 / traffic[key] - 1) < 1e-3, key
