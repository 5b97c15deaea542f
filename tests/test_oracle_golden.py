"""The oracle against the reference's own outputs (golden fixtures written
by tests/golden/make_golden.py from the reference's compiled sources).
Bit-for-bit: every flux function on 1010 state pairs, and every
fvTimeDerivative output (dU/dt, limited gradients, side states, side
fluxes) on 14 meshes / settings; where the reference is present, also live
against its compiled fvTimeDerivative on the gmsh meshes of its own cases."""
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


REF_CALC = "/root/reference/calc"


def _calc_meshes():
    out = []
    for dp, _, fs in os.walk(REF_CALC):
        for f in fs:
            p = os.path.join(dp, f)
            if f.endswith(".msh") or f.endswith(".mesh"):
                out.append(os.path.relpath(p, REF_CALC))
    return sorted(out)


CALC_MESHES = _calc_meshes() if os.path.isdir(REF_CALC) else []


@pytest.mark.parametrize("name", CALC_MESHES or ["none"])
def test_oracle_matches_reference_on_its_case_meshes(name):
    """live, on the meshes of the reference's own cases (gmsh 2 and 4 and
    EMC2 files: curved walls, stretched quads, unstructured triangles) with
    their boundary codes:
    every fvTimeDerivative output bit for bit, HLLC, 2nd order VK, Euler or
    (where a mesh has no-slip walls) Navier-Stokes"""
    from ccfd_amd import mesh as M
    from tests import cases
    if not CALC_MESHES or not B.ref_available(True):
        pytest.skip("reference cases or harness not present")
    path = os.path.join(REF_CALC, name)
    raw = M.Raw.read_emc2(path) if name.endswith(".mesh") else \
        M.Raw.read_gmsh(path)
    codes = sorted(set(raw.bcedge[:, 2].tolist()))
    ns = any(c // 100 == M.WALL for c in codes)
    bcs, items = [], [("nBC", len(codes))]
    for c in codes:
        t = c // 100
        items.append(("BCtype", c))
        if t in (M.INFLOW, M.CHARACTERISTIC):
            kw = dict(rho=1.05, mach=0.45, alpha=3.0, pressure=0.95)
            items += [(k, v) for k, v in kw.items()]
            bcs.append(M.BC(c, **kw))
        elif t == M.PRESSURE_OUT:
            items.append(("pressure", 0.9))
            bcs.append(M.BC(c, pressure=0.9))
        else:
            if t == M.WALL:
                items.append(("adiabaticWall", "T"))
            bcs.append(M.BC(c))
    mesh = M.Mesh(raw, bcs)
    prim = cases.random_state(mesh.n_elems, seed=len(name), mach=0.5)
    mu = 2e-3 if ns else 0.0
    ini = "/tmp/ccfd_calc_%d.ini" % os.getpid()
    head = [("mu", mu), ("Pr", 0.72)] if ns else []
    B.write_ini(ini, head + [("fluxFunction", 5)] + items +
                [("spatialOrder", 2), ("limiter", 2), ("venk_k", 5.0)])
    try:
        ref = B.Ref(ns=ns)
        ref.setup(ini, mesh)
        ref.set_prim(prim)
        r = ref.fvtd(0.0)
    finally:
        os.unlink(ini)
        import ctypes
        ctypes.CDLL(None).fflush(None)   # the reference's echo, captured
    cfg = B.make_cfg(flux=5, order=2, limiter=2, venk_k=5.0,
                     navierstokes=int(ns), mu=mu)
    orc = B.Oracle(mesh, cfg)
    orc.set_prim(prim)
    o = orc.fvtd(0.0, full=True)
    for k in ("u_t", "u_x", "u_y", "side", "flux"):
        assert np.array_equal(o[k], r[k]), (name, k, np.abs(o[k] - r[k]).max())


def test_exact_riemann_solution_matches_the_textbook():
    """the oracle's exact Riemann solver (src/exactRiemann.c:59-214, used by
    Godunov's flux and exact-function states) on Sod's problem against
    the published solution (Toro, Riemann Solvers and Numerical Methods for
    Fluid Dynamics, test 1: p* = 0.30313, u* = 0.92745, rho*L = 0.42632,
    rho*R = 0.26557), sampled at t = 0.2 in every region"""
    import ctypes as C
    from tests import cases
    mesh = cases.sod_mesh(10)
    cfg = B.make_cfg(left=(1.0, 0.0, 0.0, 1.0), right=(0.125, 0.0, 0.0, 0.1),
                     x0=0.5)
    L = B.oracle_lib()

    def at(x, t=0.2):
        pv = np.zeros(4)
        L.oracle_exact_func(C.byref(cfg), C.cast(mesh.ptr, C.c_void_p), 5,
                            np.array([x, 0.005]).ctypes.data_as(B.dp), t,
                            pv.ctypes.data_as(B.dp))
        return pv

    ps, us, rl, rr = 0.30313, 0.92745, 0.42632, 0.26557
    cl = (1.4 * ps / rl) ** 0.5
    t = 0.2
    # left state, fan, left star, right star, right state
    np.testing.assert_allclose(at(0.2), [1.0, 0.0, 0.0, 1.0], rtol=1e-14)
    np.testing.assert_allclose(at(0.5 + 0.5 * (us - cl) * t + 0.5 * us * t),
                               [rl, us, 0.0, ps], rtol=5e-5)
    np.testing.assert_allclose(at(0.5 + 0.5 * (us + 1.75216) * t),
                               [rr, us, 0.0, ps], rtol=5e-5)
    np.testing.assert_allclose(at(0.9), [0.125, 0.0, 0.0, 0.1], rtol=1e-14)
    # inside the rarefaction the flow is isentropic: p / rho^gamma of the
    # left state, and u + 2c / (gamma - 1) is the left Riemann invariant
    f = at(0.5 - 0.6 * t)
    c = (1.4 * f[3] / f[0]) ** 0.5
    assert 0.0 < f[1] < us and rl < f[0] < 1.0
    assert abs(f[3] / f[0] ** 1.4 - 1.0) < 1e-12
    assert abs(f[1] + 2 * c / 0.4 - 2 * 1.4 ** 0.5 / 0.4) < 1e-12
    # the shock moves at the published speed 1.75216: just behind and
    # just ahead of it
    assert at(0.5 + 1.7510 * t)[0] == pytest.approx(rr, rel=5e-5)
    assert at(0.5 + 1.7530 * t)[0] == 0.125
