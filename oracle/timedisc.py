"""Full runs of the oracle: the reference's ini reading, case setup and
timeDisc loop restated in Python over the C oracle.

TEST INFRASTRUCTURE ONLY (only tests/ use it).

  Ini            readInTools semantics (src/readInTools.c:38-405): keys
                 lower-cased, values keep [A-Za-z0-9_.,-], repeated keys are
                 consumed in order
  setup_case     initEquation / initBoundary / initMesh / initInitialCondition
                 / initFV / initTimeDisc (src/main.c:101-121) for Cartesian and
                 gmsh meshes, icType 1 and 2
  run_case       timeDisc (src/timeDiscretization.c:559-777) with CSV output
                 (csvOutput, src/output.c:56-113)

Pinned by the reference's own make check outputs (tests/test_make_check_oracle.py)
and, for Ini, by the reference's compiled readInTools on every case file
(tests/test_ini_host.py).
"""
import math
import os

import numpy as np

from ccfd_amd import mesh as M
from oracle import bindings as B


class Ini:
    def __init__(self, path):
        self.cmds = []
        with open(path) as f:
            for line in f:
                key = ""
                i = 0
                while i < len(line):
                    ch = line[i]
                    if ch.isalnum() or ch == "_":
                        key += ch.lower()
                    elif ch in "!#=":
                        break
                    i += 1
                if not key:
                    continue
                val = ""
                while i < len(line):
                    ch = line[i]
                    if ch.isalnum() or ch in "_.,-":
                        val += ch
                    elif ch in "!#":
                        break
                    i += 1
                self.cmds.append([key, val])

    def _find(self, key, default):
        k = key.lower()
        for n, (kk, v) in enumerate(self.cmds):
            if kk == k:
                del self.cmds[n]
                return v
        if default is None:
            raise KeyError("ini file missing necessary keyword: " + k)
        return default

    def count(self, key):
        return sum(1 for kk, _ in self.cmds if kk == key.lower())

    def str(self, key, default=None):
        return self._find(key, default)

    def int(self, key, default=None):
        return int(float(self._find(key, default)))

    def dbl(self, key, default=None):
        return float(self._find(key, default))

    def bool(self, key, default=None):
        return self._find(key, default) in ("T", "True", "t", "true")

    def arr(self, key, n, default=None, conv=float):
        v = self._find(key, default).split(",")
        return [conv(float(x)) for x in v[:n]]


def setup_case(ini_path):
    """parse a reference ini and build mesh, oracle configuration and
    initial state"""
    ini = Ini(ini_path)
    d = os.path.dirname(os.path.abspath(ini_path))
    c = {}
    c["stationary"] = ini.bool("stationary", "T")
    c["fileName"] = ini.str("fileName")
    c["IOtimeInterval"] = ini.dbl("IOtimeInterval")
    c["IOiterInterval"] = int(ini.dbl("IOiterInterval"))
    c["outputFormat"] = ini.int("outputFormat", "1")
    # initEquation
    gam = ini.dbl("gamma", "1.4")
    ini.dbl("R", "287")
    c["flux"] = ini.int("fluxFunction")
    ini.bool("calcSource", "F")
    # initBoundary
    nBC = ini.int("nBC")
    bcs = []
    for _ in range(nBC):
        code = ini.int("BCtype")
        t = code // 100
        if t in (M.INFLOW, M.CHARACTERISTIC):
            rho = ini.dbl("rho"); mach = ini.dbl("mach")
            alpha = ini.dbl("alpha"); p = ini.dbl("pressure")
            bcs.append(M.BC(code, rho=rho, mach=mach, alpha=alpha,
                            pressure=p, gamma=gam))
        elif t == M.PRESSURE_OUT:
            bcs.append(M.BC(code, pressure=ini.dbl("pressure")))
        elif t == M.PERIODIC:
            bcs.append(M.BC(code, connection=tuple(ini.arr("connection", 2))))
        elif t == M.EXACTSOL:
            bcs.append(M.BC(code, exact_func=ini.int("BCexactFunc")))
        else:
            bcs.append(M.BC(code))
    # initMesh
    mtype = ini.int("meshType", "1")
    if mtype == 1:
        iMax, jMax = ini.int("nElemsX"), ini.int("nElemsY")
        x0 = ini.arr("X0", 2)
        x1 = ini.arr("Xmax", 2)
        nseg = ini.arr("nBCsegments", 4, conv=int)
        side = []
        for s in range(4):
            if nseg[s] != 1:
                raise NotImplementedError("multi-segment Cartesian sides")
            side.append(ini.int("meshBCtype"))
        raw = M.cartesian(iMax, jMax, tuple(x0), tuple(x1), tuple(side))
    else:
        fmt = ini.str("meshFormat")
        fn = ini.str("meshFile")
        raw = M.Raw.read_gmsh(os.path.join(d, fn + fmt))
    mesh = M.Mesh(raw, bcs)
    # initInitialCondition
    ic = ini.int("icType")
    left = right = (1.0, 0.0, 0.0, 1.0)
    x0i = 0.5
    prim = np.zeros((mesh.n_elems, 4))
    if ic == 2:
        ef = ini.int("exactFunc")
        if ef != 5:
            raise NotImplementedError("exact function %d initial state" % ef)
        x0i = ini.dbl("RP_1D_interface", "0.5")
        left = tuple(ini.arr("StateLeft", 4))
        right = tuple(ini.arr("StateRight", 4))
        x = mesh.bary[:, 0]
        prim[x <= x0i] = left
        prim[x > x0i] = right
        prim[:, 2] = 0.0
    elif ic == 1:
        nd = ini.int("nDomains", "1")
        if nd != 1:
            raise NotImplementedError("multi-domain initial state")
        ini.int("domainID")
        rho = ini.dbl("rho"); mach = ini.dbl("mach")
        alpha = ini.dbl("alpha"); p = ini.dbl("pressure")
        cc = math.sqrt(gam * p / rho) * mach
        pi = math.acos(-1.0)
        prim[:] = (rho, cc * math.cos(alpha * pi / 180.0),
                   cc * math.sin(alpha * pi / 180.0), p)
    # initFV
    order = ini.int("spatialOrder", "1")
    lim, vk = 1, 1.0
    if order == 2:
        lim = ini.int("limiter", "1")
        if lim == 2:
            vk = ini.dbl("venk_k", "1")
    # initTimeDisc
    c["cfl"] = ini.dbl("CFL", "0.9")
    c["dfl"] = ini.dbl("DFL", "0.9")
    c["ts1d"] = ini.bool("timeStep1D", "F")
    c["timeOrder"] = ini.int("timeOrder", "1")
    c["implicit"] = ini.bool("implicit", "F")
    c["nRK"] = ini.int("nRKstages", "1")
    c["maxIter"] = ini.int("maxIter", "100000")
    c["tEnd"] = ini.dbl("tEnd")
    if c["implicit"]:
        c["nKdim"] = ini.int("nKdim", "5")
        c["epsNewton"] = ini.dbl("epsNewton", "0.001")
        c["epsGMRES"] = ini.dbl("epsGMRES", "0.001")
        c["nNewtonIter"] = ini.int("nNewtonIter", "20")
        c["gammaEW"] = ini.dbl("gammaEW", "0.9")
        c["precond"] = ini.bool("precond", "F")
    cfg = B.make_cfg(gamma=gam, flux=c["flux"], order=order, limiter=lim,
                     venk_k=vk, left=left, right=right, x0=x0i,
                     cfl=c["cfl"], dfl=c["dfl"], ts1d=int(c["ts1d"]),
                     nrk=c["nRK"], torder=c["timeOrder"],
                     dxref=math.sqrt(1.0 / (mesh.total_area_q * mesh.n_elems)))
    c.update(gam=gam, order=order, limiter=lim, venk_k=vk, left=left,
             right=right, x0=x0i)
    return mesh, cfg, prim, c


def csv_rows(mesh, prim):
    """csvOutput (src/output.c:56-113): rows (x, rho, u, p) in element order,
    bubble-sorted by x (stable), printed with %15.9f"""
    x = mesh.bary[:, 0]
    order = np.argsort(x, kind="stable")
    rows = np.stack([x, prim[:, 0], prim[:, 1], prim[:, 3]], 1)[order]
    text = "CoordinateX, Density, Velocity, Pressure\n" + "".join(
        "%15.9f,%15.9f,%15.9f,%15.9f\n" % tuple(r) for r in rows)
    return rows, text


def run_case(ini_path, max_iter=None):
    """timeDisc of the oracle; returns (final time, iterations, final prim,
    mesh, case dict)"""
    mesh, cfg, prim, c = setup_case(ini_path)
    orc = B.Oracle(mesh, cfg)
    orc.set_prim(prim)
    t = 0.0
    printTime = (math.floor(t / c["IOtimeInterval"]) + 1) * c["IOtimeInterval"]
    printIter = (0 // c["IOiterInterval"] + 1) * c["IOiterInterval"]
    tEnd = c["tEnd"]
    maxIter = max_iter or c["maxIter"]
    ic = None
    if c["implicit"]:
        ic = B.ImplicitCfg(c["nKdim"], c["nNewtonIter"], int(c["precond"]),
                           c["epsNewton"], c["epsGMRES"], c["gammaEW"])
    it = 0
    for it in range(1, maxIter + 1):
        dt, _, nonfinite = orc.calc_dt(t, printTime, tEnd)
        if nonfinite:
            raise FloatingPointError("convective time step NaN")
        if ic is None:
            orc.explicit_step(t, dt)
        else:
            rc, _, _, _ = orc.implicit_step(ic, t, dt)
            if rc:
                raise RuntimeError("implicit step failed: %d" % rc)
        t += dt
        if tEnd - t <= 1e-15:
            break
        if (printTime - t <= 1e-15) or (it == printIter):
            if printTime - t <= 1e-15:
                printTime += c["IOtimeInterval"]
            if it == printIter:
                printIter += c["IOiterInterval"]
    return t, it, orc.prim(), mesh, c


def check_metric(rows, target_csv):
    """check/check.py:26-35: L2 over columns x, density, velocity"""
    tgt = np.genfromtxt(target_csv, skip_header=1, delimiter=",")[:, 0:3]
    diff = np.sqrt(((tgt - rows[:, 0:3]) ** 2).sum(axis=0))
    return float(np.sqrt((diff ** 2).sum()))
