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
    entry must be the scaled FETCH_SIZE + WRITE_SIZE bytes per dispatch
    that the committed rocprofv3 summary of that kernel lists"""
    import json
    import re
    prof = os.path.join(ROOT, "profiles")
    with open(os.path.join(prof, "pmc_traffic.json")) as f:
        traffic = json.load(f)["kernels"]
    assert set(traffic) == {"fused_face_fp64", "fused_face_fp32"}
    for key, ent in traffic.items():
        text = open(os.path.join(prof, ent["summary"])).read()
        head = text.index("## HBM traffic")
        tab = text[head:].splitlines()

        def last_cell(name):
            line = [x for x in tab if x.startswith("| " + name)][0]
            return float(re.findall(r"[0-9.]+e[0-9]+", line)[-1])
        fetch, write, total = (last_cell("FETCH_SIZE"), last_cell("WRITE_SIZE"),
                               last_cell("total"))
        assert abs(ent["bytes_per_launch"] / total - 1) < 1e-3, key
        assert abs((fetch + write) / total - 1) < 2e-3, key


def test_traffic_figure_only_for_the_sources_it_was_measured_on(tmp_path,
                                                                monkeypatch):
    """the committed counter figure is used only while the device sources
    hash to what it was measured on; otherwise the line says it is stale"""
    sys.path.insert(0, ROOT)
    import json

    import bench
    (tmp_path / "profiles").mkdir()
    d = {"kernel_source_sha": "aaaa", "kernels": {"fused_face_fp64": {
        "bytes_per_launch": 1.5e9, "summary": "x.md", "command": "cmd"}}}
    (tmp_path / "profiles" / "pmc_traffic.json").write_text(json.dumps(d))
    monkeypatch.setattr(bench, "ROOT", str(tmp_path))
    monkeypatch.setattr(bench, "kernel_source_sha", lambda: "aaaa")
    t, src = bench.pmc_traffic("fused_face_fp64")
    assert t == 1.5e9 and "x.md" in src
    assert bench.pmc_traffic("fused_face_fp32")[0] is None
    monkeypatch.setattr(bench, "kernel_source_sha", lambda: "bbbb")
    t, src = bench.pmc_traffic("fused_face_fp64")
    assert t is None and "stale" in src


def test_committed_traffic_figures_match_the_sources():
    """profiles/pmc_traffic.json names the hash of the committed device
    sources"""
    sys.path.insert(0, ROOT)
    import json

    import bench
    with open(os.path.join(ROOT, "profiles", "pmc_traffic.json")) as f:
        d = json.load(f)
    assert d["kernel_source_sha"] == bench.kernel_source_sha()
    for k, e in d["kernels"].items():
        assert os.path.exists(os.path.join(ROOT, "profiles", e["summary"])), k


def test_ut_parity_summary():
    sys.path.insert(0, ROOT)
    import bench
    import numpy as np
    rng = np.random.default_rng(0)
    ref = rng.normal(size=(50, 4))
    p = bench._ut_parity(ref.copy(), ref)
    assert p["bitwise"] and p["max_rel_err"] == 0.0
    got = ref.copy()
    got[7, 3] += 1e-9 * np.abs(ref[:, 3]).max()
    p = bench._ut_parity(got, ref)
    assert not p["bitwise"]
    assert abs(p["max_rel_err"] - 1e-9) < 1e-12
    assert p["density_ut_rel_l2"] == 0.0 and p["energy_ut_rel_linf"] > 0.0
