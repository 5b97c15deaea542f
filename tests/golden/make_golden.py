#!/usr/bin/env python3
"""Write the golden fixtures of the reference's operator.

Runs the reference's own fvTimeDerivative / convectiveFlux, compiled from
/root/reference/src by oracle/Makefile into oracle/_ref/ (see
oracle/ref_harness.c), on small seeded cases and stores inputs and outputs
as .npz vectors under tests/golden/. Also copies the reference's make-check
inputs and targets (check/*.ini, check/targetOutput/*.csv; data files) into
tests/golden/check/.

Only runs where /root/reference exists; the fixtures are committed so the
tests need neither the reference nor this script.

    python tests/golden/make_golden.py
"""
import glob
import hashlib
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from ccfd_amd import mesh as M          # noqa: E402
from oracle import bindings as B       # noqa: E402
from tests import cases                # noqa: E402

REF = "/root/reference"


def ref_sources_digest():
    h = hashlib.sha256()
    for f in sorted(glob.glob(os.path.join(REF, "src", "*.[ch]"))):
        with open(f, "rb") as fh:
            h.update(os.path.basename(f).encode())
            h.update(fh.read())
    return h.hexdigest()


def bc_ini(bcs):
    items = [("nBC", len(bcs))]
    for b in bcs:
        items.append(("BCtype", b.code))
        if b.type in (M.INFLOW, M.CHARACTERISTIC):
            items += [("rho", b.kw["rho"]), ("mach", b.kw["mach"]),
                      ("alpha", b.kw["alpha"]), ("pressure", b.kw["pressure"])]
        elif b.type == M.PRESSURE_OUT:
            items.append(("pressure", b.kw["pressure"]))
        elif b.type == M.PERIODIC:
            items.append(("connection", b.connection))
        elif b.type == M.EXACTSOL:
            items.append(("BCexactFunc", b.exact_func))
        elif b.type == M.WALL:
            items.append(("adiabaticWall", "T"))
    return items


class BCspec(M.BC):
    def __init__(self, code, **kw):
        super().__init__(code, **kw)
        self.kw = kw


def run_case(name, raw, bcs, prim, flux, order, limiter=1, venk_k=1.0,
             ns=False, mu=0.0, time=0.0, riemann=None, out=None):
    mesh = M.Mesh(raw, bcs)
    ini = "/tmp/ccfd_golden_%d.ini" % os.getpid()
    items = []
    if ns:
        items += [("mu", mu), ("Pr", 0.72)]
    items += [("fluxFunction", flux)] + bc_ini(bcs)
    items += [("spatialOrder", order), ("limiter", limiter),
              ("venk_k", venk_k)]
    B.write_ini(ini, items)
    ref = B.Ref(ns=ns)
    if riemann is not None:
        ref.setup(ini, mesh, left=riemann[0], right=riemann[1], x0=riemann[2])
    else:
        ref.setup(ini, mesh)
    ref.set_prim(prim)
    r = ref.fvtd(time)
    os.unlink(ini)
    bc_rows = np.array([[b.code, b.exact_func] + list(b.pvar) +
                        list(b.connection) for b in bcs], dtype=np.float64)
    d = dict(xy=raw.xy, tri=raw.tri, quad=raw.quad, bcedge=raw.bcedge,
             bcs=bc_rows, prim=prim,
             cfg=np.array([flux, order, limiter, venk_k, 1 if ns else 0, mu,
                           time], dtype=np.float64),
             u_t=r["u_t"], u_x=r["u_x"], u_y=r["u_y"], side=r["side"],
             flux=r["flux"])
    if riemann is not None:
        d["riemann"] = np.array(list(riemann[0]) + list(riemann[1]) +
                                [riemann[2]], dtype=np.float64)
    out[name] = d
    print("case %-28s %6d elements" % (name, mesh.n_elems))


def flux_states(n, seed):
    """random rotated states over wide density / pressure / Mach ranges plus
    the special cases each flux branches on"""
    rng = np.random.Generator(np.random.PCG64(seed))

    def rnd(k):
        s = np.empty((k, 4))
        s[:, 0] = np.exp(rng.uniform(-3, 3, k))
        s[:, 3] = np.exp(rng.uniform(-3, 3, k))
        c = np.sqrt(1.4 * s[:, 3] / s[:, 0])
        s[:, 1] = c * rng.uniform(-3, 3, k)
        s[:, 2] = rng.normal(size=k)
        return s
    L, R = rnd(n), rnd(n)
    special = [
        # supersonic to the right / left
        ((1.0, 3.0, 0.1, 1.0), (0.5, 3.2, -0.2, 0.7)),
        ((1.0, -3.0, 0.1, 1.0), (0.5, -3.2, -0.2, 0.7)),
        # sonic points (entropy fix)
        ((1.0, 1.1832159566199232, 0.0, 1.0), (0.9, 1.0, 0.0, 0.8)),
        ((1.0, -1.1832159566199232, 0.0, 1.0), (1.1, -1.3, 0.0, 1.2)),
        # stationary contact
        ((1.0, 0.0, 0.3, 1.0), (0.125, 0.0, -0.4, 1.0)),
        # strong shock, pressure ratio 1e3
        ((1.0, 0.0, 0.0, 1000.0), (1.0, 0.0, 0.0, 1.0)),
        # near vacuum
        ((1e-6, 0.0, 0.0, 1e-6), (1.0, 0.0, 0.0, 1.0)),
        # identical states (consistency)
        ((1.3, 0.4, -0.2, 2.0), (1.3, 0.4, -0.2, 2.0)),
        ((0.5, -2.5, 1.0, 0.3), (0.5, -2.5, 1.0, 0.3)),
        # Sod
        ((1.0, 0.0, 0.0, 1.0), (0.125, 0.0, 0.0, 0.1)),
    ]
    Ls = np.array([a for a, _ in special])
    Rs = np.array([b for _, b in special])
    return np.vstack([L, Ls]), np.vstack([R, Rs])


def main():
    if not os.path.isdir(os.path.join(REF, "src")):
        sys.exit("reference tree not found at " + REF)
    if not (B.ref_available(False) and B.ref_available(True)):
        sys.exit("build the harness first: make -C oracle ref")
    out = {}

    # F1: every flux function on random + special states
    L, R = flux_states(1000, 7)
    ref = B.Ref()
    fx = {"L": L, "R": R}
    for code in range(1, 12):
        fx["F%02d" % code] = ref.riemann(code, L, R)
    np.savez_compressed(os.path.join(HERE, "riemann_flux.npz"), **fx)
    print("riemann_flux.npz: %d state pairs x 11 fluxes" % len(L))

    # F2: operator on meshes / states / settings
    sod = M.cartesian(100, 1, (0.0, 0.0), (1.0, 0.01), (101, 401, 101, 401))
    sod_bcs = [BCspec(101), BCspec(401)]
    sodm = M.Mesh(sod, sod_bcs)
    p_sod = cases.sod_state(sodm)
    p_rnd = cases.random_state(sodm.n_elems, seed=3)
    run_case("sod100_o1_hll", sod, sod_bcs, p_sod, 3, 1, out=out)
    run_case("sod100_o2_bj_hll", sod, sod_bcs, p_rnd, 3, 2, 1, out=out)
    run_case("sod100_o2_vk10_hll", sod, sod_bcs, p_rnd, 3, 2, 2, 10.0, out=out)
    c1 = M.cartesian(200, 2, (0.0, 0.0), (1.0, 0.01), (101, 401, 101, 401))
    c1m = M.Mesh(c1, sod_bcs)
    run_case("sod200x2_o2_bj_hllc", c1, sod_bcs,
             cases.random_state(c1m.n_elems, seed=4), 5, 2, 1, out=out)

    raw, pbcs = M.jittered_delaunay(30, 0.3, 2, periodic=True)
    pbcs = [BCspec(701, connection=(0.0, 1.0)), BCspec(702),
            BCspec(711, connection=(1.0, 0.0)), BCspec(712)]
    dm = M.Mesh(raw, pbcs)
    p_sm = M.smooth_state(dm.bary, seed=2)
    for code in (2, 3, 5, 9):
        run_case("delaunay30_vk5_f%d" % code, raw, pbcs, p_sm, code, 2, 2,
                 5.0, out=out)

    mix = M.cartesian(24, 12, (0.0, 0.0), (2.0, 1.0), (101, 801, 301, 501))
    mbcs = [BCspec(101), BCspec(801, pressure=0.9),
            BCspec(301, rho=1.1, mach=0.6, alpha=10.0, pressure=1.05),
            BCspec(501, rho=0.95, mach=0.4, alpha=-5.0, pressure=1.0)]
    mm = M.Mesh(mix, mbcs)
    p_mx = cases.random_state(mm.n_elems, seed=11, mach=0.5)
    run_case("mixedbc_o1_hllc", mix, mbcs, p_mx, 5, 1, out=out)
    run_case("mixedbc_o2_vk5_roe", mix, mbcs, p_mx, 2, 2, 2, 5.0, out=out)
    run_case("mixedbc_o2_bj_ausmd", mix, mbcs, p_mx, 9, 2, 1, out=out)

    # mixed triangles and quadrilaterals (absent slots in the device layout)
    mixq = cases.split_quads(mix, seed=3)
    mqm = M.Mesh(mixq, mbcs)
    p_mq = cases.random_state(mqm.n_elems, seed=13, mach=0.5)
    run_case("mixedmesh_o2_vk5_hllc", mixq, mbcs, p_mq, 5, 2, 2, 5.0, out=out)
    run_case("mixedmesh_o1_roe", mixq, mbcs, p_mq, 2, 1, out=out)

    # exact-function boundaries (Sod Riemann problem at t = 0.1)
    ex = M.cartesian(50, 1, (0.0, 0.0), (1.0, 0.02), (101, 601, 101, 601))
    ebcs = [BCspec(101), BCspec(601, exact_func=5)]
    em = M.Mesh(ex, ebcs)
    run_case("exactbc_sod_t01_o2_hll", ex, ebcs,
             cases.random_state(em.n_elems, seed=5, mach=0.3), 3, 2, 1,
             time=0.1, riemann=(cases.SOD_LEFT, cases.SOD_RIGHT, 0.5),
             out=out)

    # Navier-Stokes build: no-slip walls, far field
    nsr = M.cartesian(20, 10, (0.0, 0.0), (1.0, 0.5), (201, 501, 101, 401))
    nbcs = [BCspec(201), BCspec(501, rho=1.0, mach=0.3, alpha=0.0,
                                 pressure=1.0), BCspec(101), BCspec(401)]
    nm = M.Mesh(nsr, nbcs)
    p_ns = cases.random_state(nm.n_elems, seed=5, mach=0.3)
    for order in (1, 2):
        run_case("ns_o%d_roe" % order, nsr, nbcs, p_ns, 2, order, 2, 5.0,
                 ns=True, mu=0.01, out=out)

    flat = {}
    for name, d in out.items():
        for k, v in d.items():
            flat[name + "/" + k] = v
    flat["_ref_sources_sha256"] = np.array(ref_sources_digest())
    np.savez_compressed(os.path.join(HERE, "operator_cases.npz"), **flat)
    print("operator_cases.npz: %d cases" % len(out))

    # make check inputs and targets (data)
    dst = os.path.join(HERE, "check")
    os.makedirs(os.path.join(dst, "targetOutput"), exist_ok=True)
    for f in glob.glob(os.path.join(REF, "check", "*ini")):
        shutil.copy(f, dst)
    for f in glob.glob(os.path.join(REF, "check", "targetOutput", "*.csv")):
        shutil.copy(f, os.path.join(dst, "targetOutput"))
    print("check/: %d ini, %d targets" % (
        len(glob.glob(os.path.join(dst, "*ini"))),
        len(glob.glob(os.path.join(dst, "targetOutput", "*.csv")))))


if __name__ == "__main__":
    devnull = os.open(os.devnull, os.O_WRONLY)
    saved = os.dup(1)
    os.dup2(devnull, 1)        # the reference echoes every ini key
    try:
        sys.stdout = os.fdopen(saved, "w")
        main()
    finally:
        os.dup2(saved, 1)
