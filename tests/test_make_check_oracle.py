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
