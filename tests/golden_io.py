"""Load the committed golden fixtures (tests/golden/*.npz) into meshes,
states and expected outputs."""
import os

import numpy as np

from ccfd_amd import mesh as M

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def flux_vectors():
    return dict(np.load(os.path.join(GOLDEN, "riemann_flux.npz")))


def operator_cases():
    z = np.load(os.path.join(GOLDEN, "operator_cases.npz"))
    cases = {}
    for key in z.files:
        if "/" not in key:
            continue
        name, field = key.split("/", 1)
        cases.setdefault(name, {})[field] = z[key]
    return cases


def build(case):
    """(mesh, bcs, prim, settings dict) of one operator case"""
    bcs = []
    for row in case["bcs"]:
        code, ef = int(row[0]), int(row[1])
        bcs.append(M.BC(code, pvar=list(row[2:6]), connection=tuple(row[6:8]),
                        exact_func=ef))
    raw = M.Raw(case["xy"], case["tri"], case["quad"], case["bcedge"])
    mesh = M.Mesh(raw, bcs)
    c = case["cfg"]
    cfg = dict(flux=int(c[0]), order=int(c[1]), limiter=int(c[2]),
               venk_k=float(c[3]), ns=int(c[4]), mu=float(c[5]),
               time=float(c[6]))
    if "riemann" in case:
        r = case["riemann"]
        cfg["left"] = tuple(r[0:4])
        cfg["right"] = tuple(r[4:8])
        cfg["x0"] = float(r[8])
    return mesh, bcs, case["prim"], cfg
