"""ccfd's own regression suite (make check: check/check.py over the 14 active
check/*.ini Sod cases, Makefile:116-117) reproduced by the oracle on the CPU.
Targets are the reference's stored outputs (check/targetOutput/*.csv); the
metric is check.py's combined L2 over x, density and velocity, 'ok' below
1e-5. The printed CSV is compared, as check.py reads it."""
import glob
import os

import numpy as np
import pytest

from oracle import timedisc as TD

HERE = os.path.dirname(os.path.abspath(__file__))
CHECK = os.path.join(HERE, "golden", "check")
INIS = sorted(glob.glob(os.path.join(CHECK, "*.ini")))


def target(name):
    return sorted(glob.glob(os.path.join(CHECK, "targetOutput", name + "_*")))[-1]


@pytest.mark.parametrize("ini", INIS, ids=[os.path.basename(i) for i in INIS])
def test_oracle_reproduces_make_check(ini):
    name = os.path.basename(ini)[:-4]
    t, it, prim, mesh, c = TD.run_case(ini)
    assert abs(t - c["tEnd"]) < 1e-12
    rows, text = TD.csv_rows(mesh, prim)
    printed = np.genfromtxt(text.splitlines(), skip_header=1, delimiter=",")
    metric = TD.check_metric(printed, target(name))
    print("%s: %d iterations, check metric %.3e" % (name, it, metric))
    assert metric < 1e-5


def test_fourteen_active_cases():
    assert len(INIS) == 14


def test_time_step_closed_forms():
    """calcTimeStep (src/timeDiscretization.c:249-329) on a uniform
    Cartesian grid, where it reduces to the textbook limits: convective
    CFL / ((|u| + c) / dx + (|v| + c) / dy), viscous
    DFL rho / (4 max(4/3, gamma/Pr) (mu/rho) (1/dx^2 + 1/dy^2)) with the
    flag set when it governs, the 1-D form CFL dx / (|u| + c), and the
    clamp to the next output time"""
    import numpy as np
    from ccfd_amd import mesh as M
    from oracle import bindings as B
    ni, nj = 40, 5
    dx, dy = 1.0 / ni, 0.01 / nj
    raw = M.cartesian(ni, nj, (0.0, 0.0), (1.0, 0.01), (101, 401, 101, 401))
    mesh = M.Mesh(raw, [M.BC(101), M.BC(401)])
    rho, u, v, p, gam = 1.3, 0.4, -0.25, 0.9, 1.4
    prim = np.tile([rho, u, v, p], (mesh.n_elems, 1))
    c = (gam * p / rho) ** 0.5
    conv = 0.9 / ((abs(u) + c) / dx + (abs(v) + c) / dy)

    def dt_of(**kw):
        o = B.Oracle(mesh, B.make_cfg(**kw))
        o.set_prim(prim)
        return o, o.calc_dt(0.0, 1e9, 1e9)

    _, (dt, visc, nf) = dt_of(cfl=0.9)
    assert abs(dt / conv - 1) < 1e-13 and not visc and not nf
    flags = set()
    for mu, Pr in ((1e-7, 0.72), (1e-4, 0.72), (1e-3, 0.72), (1e-3, 2.0)):
        vdt = 0.7 * rho / (4 * max(4 / 3, gam / Pr) * (mu / rho) *
                           (1 / dx ** 2 + 1 / dy ** 2))
        _, (dt, visc, _) = dt_of(cfl=0.9, dfl=0.7, mu=mu, Pr=Pr,
                                 navierstokes=1)
        assert abs(dt / min(conv, vdt) - 1) < 1e-13, (mu, Pr)
        assert visc == (vdt < conv)
        flags.add(visc)
    assert flags == {True, False}      # both limits govern somewhere
    _, (dt, _, _) = dt_of(cfl=0.9, ts1d=1)
    assert abs(dt / (0.9 * dx / (abs(u) + c)) - 1) < 1e-13
    # output clamp: a step that would pass the output time ends on it, one
    # that would come within half a step of it takes half the remainder
    o, _ = dt_of(cfl=0.9)
    assert o.calc_dt(0.0, 0.5 * conv, 1e9)[0] == 0.5 * conv
    assert abs(o.calc_dt(0.0, 1.2 * conv, 1e9)[0] - 0.6 * conv) < 1e-18
    assert o.calc_dt(0.0, 1e9, 0.25 * conv)[0] == 0.25 * conv


def test_rk_tables_are_the_reference_tables():
    """every (stages, order) coefficient table of the oracle equals the one
    in the reference's source (src/timeDiscretization.c:94-178), read as
    text: make check runs only the order-2 tables and forward Euler"""
    import re
    from oracle import bindings as B
    src = "/root/reference/src/timeDiscretization.c"
    if not os.path.exists(src):
        pytest.skip("reference sources not present")
    lines = open(src).read().splitlines()
    start = next(i for i, x in enumerate(lines) if "nRKstages == 1" in x)
    tables, stages, order = {(1, 1): {1: 1.0}}, None, None
    for i in range(start, len(lines)):
        x = lines[i]
        if "default:" in x:
            break
        m = re.search(r"case (\d+):", x)
        if m:
            if "switch (timeOrder)" in lines[i + 1]:
                stages = int(m.group(1))
            else:
                order = int(m.group(1))
            continue
        m = re.search(r"RKcoeff\[(\d+)\] = ([0-9.]+);", x)
        if m and stages:
            tables.setdefault((stages, order), {})[int(m.group(1))] = \
                float(m.group(2))
    assert len(tables) == 10
    for (s, o), coeff in tables.items():
        got = B.rk_coeffs(s, o)
        assert got is not None, (s, o)
        assert sorted(coeff) == list(range(1, s + 1))
        for k, v in coeff.items():
            assert got[k] == v, (s, o, k)
        assert got[0] == 0.0
    assert B.rk_coeffs(2, 2) is None


def test_global_residual_is_the_area_weighted_rms():
    """globalResidual (src/analyze.c:521-539): per variable the
    area-weighted root mean square of dU/dt over the domain; a constant
    field gives its magnitude"""
    import numpy as np
    from oracle import bindings as B
    from tests import cases
    mesh = cases.delaunay_mesh(14, seed=7, periodic=False)
    o = B.Oracle(mesh, B.make_cfg())
    rng = np.random.default_rng(3)
    ut = rng.normal(size=(mesh.n_elems, 4)) * [1.0, 10.0, 0.1, 3.0]
    want = np.sqrt(np.average(ut ** 2, axis=0, weights=mesh.area))
    np.testing.assert_allclose(o.residual(ut), want, rtol=1e-13)
    c = np.array([-0.5, 2.0, 0.0, 7.0])
    np.testing.assert_allclose(o.residual(np.tile(c, (mesh.n_elems, 1))),
                               np.abs(c), rtol=1e-13, atol=0)
