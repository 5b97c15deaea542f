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
