"""The oracle against the reference's own outputs (golden fixtures written
by tests/golden/make_golden.py from the reference's compiled sources).
Bit-for-bit: every flux function on 1010 state pairs, and every
fvTimeDerivative output (dU/dt, limited gradients, side states, side
fluxes) on 14 meshes / settings."""
import os

import numpy as np
import pytest

from oracle import bindings as B
from tests import golden_io

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CASES = golden_io.operator_cases()


@pytest.mark.parametrize("code", list(range(1, 12)))
def test_oracle_riemann_matches_reference(code):
    fx = golden_io.flux_vectors()
    F = B.riemann(B.make_cfg(), code, fx["L"], fx["R"])
    assert np.array_equal(F, fx["F%02d" % code], equal_nan=True)


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_operator_matches_reference(name):
    mesh, bcs, prim, c = golden_io.build(CASES[name])
    cfg = B.make_cfg(flux=c["flux"], order=c["order"], limiter=c["limiter"],
                     venk_k=c["venk_k"], navierstokes=c["ns"], mu=c["mu"],
                     left=c.get("left", (1, 0, 0, 1)),
                     right=c.get("right", (0.125, 0, 0, 0.1)),
                     x0=c.get("x0", 0.5))
    orc = B.Oracle(mesh, cfg)
    orc.set_prim(prim)
    o = orc.fvtd(c["time"], full=True)
    g = CASES[name]
    for k in ("u_t", "u_x", "u_y", "side", "flux"):
        assert np.array_equal(o[k], g[k]), (name, k, np.abs(o[k] - g[k]).max())


def test_fixture_sanity():
    """the fixtures exercise what they claim: supersonic and subsonic
    branches, all BC kinds, non-trivial limiting"""
    fx = golden_io.flux_vectors()
    c = np.sqrt(1.4 * fx["L"][:, 3] / fx["L"][:, 0])
    mach = fx["L"][:, 1] / c
    assert (mach > 1).sum() > 50 and (mach < -1).sum() > 50
    assert (np.abs(mach) < 1).sum() > 50
    g = CASES["sod100_o2_vk10_hll"]
    assert np.abs(g["u_x"]).max() > 0
    codes = set()
    for case in CASES.values():
        codes |= set(int(r[0]) // 100 for r in case["bcs"])
    assert {1, 2, 3, 4, 5, 6, 7, 8} <= codes


def test_reference_harness_keeps_its_own_fvtimederivative():
    """The drop-in library exports fvTimeDerivative too; with both loaded in
    one process (as bench.py's CPU baseline does) the reference harness must
    still call its own. Run in a child process: the drop-in's version would
    exit(1)."""
    import subprocess
    import sys
    from oracle import bindings as B
    if not B.ref_available(False):
        pytest.skip("reference harness not built")
    code = (
        "import sys; sys.path.insert(0, %r)\n"
        "from ccfd_amd import mesh as M\n"
        "from oracle import bindings as B\n"
        "raw, bcs = M.jittered_delaunay(12, 0.3, 2, periodic=True)\n"
        "mesh = M.Mesh(raw, bcs)\n"
        "B.write_ini('/tmp/ccfd_sym_%%d.ini' %% __import__('os').getpid(), "
        "[('fluxFunction', 5), ('nBC', 4), ('BCtype', 701), "
        "('connection', (0.0, 1.0)), ('BCtype', 702), "
        "('connection', (0.0, 0.0)), ('BCtype', 711), "
        "('connection', (1.0, 0.0)), ('BCtype', 712), "
        "('connection', (0.0, 0.0)), ('spatialOrder', 2), ('limiter', 2), "
        "('venk_k', 5.0)])\n"
        "ref = B.Ref(omp=True)\n"
        "ref.setup('/tmp/ccfd_sym_%%d.ini' %% __import__('os').getpid(), mesh)\n"
        "ref.set_prim(M.smooth_state(mesh.bary, seed=2))\n"
        "ref.fvtd_only(0.0)\n"
        "print('REF-OK')\n" % ROOT)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0 and "REF-OK" in r.stdout, r.stdout[-500:]
