"""make check on the GPU path: the drop-in binary (`ccfd case.ini`, C host
driver over the HIP kernels) runs the reference's 14 active Sod cases and
its CSV output is compared the way check/check.py does (L2 over x, density,
velocity < 1e-5), and against the oracle's run (same printed digits)."""
import glob
import os
import subprocess
import tempfile

import numpy as np
import pytest

from oracle import timedisc as TD

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CHECK = os.path.join(HERE, "golden", "check")
INIS = sorted(glob.glob(os.path.join(CHECK, "*.ini")))
CCFD = os.path.join(ROOT, "ccfd_amd", "lib", "ccfd")


@pytest.mark.parametrize("ini", INIS, ids=[os.path.basename(i) for i in INIS])
def test_dropin_reproduces_make_check(ini):
    name = os.path.basename(ini)[:-4]
    with tempfile.TemporaryDirectory() as d:
        r = subprocess.run([CCFD, ini], cwd=d, stdout=subprocess.PIPE,
                           stderr=subprocess.STDOUT, text=True, timeout=120)
        assert r.returncode == 0, r.stdout[-2000:]
        out = sorted(glob.glob(os.path.join(d, name + "_*.csv")))[-1]
        got = np.genfromtxt(out, skip_header=1, delimiter=",")
        text = open(out).read()
    tgt = sorted(glob.glob(os.path.join(CHECK, "targetOutput", name + "_*")))[-1]
    metric = TD.check_metric(got, tgt)
    print("%s: check metric %.3e" % (name, metric))
    assert metric < 1e-5
    # the oracle's run prints the same digits
    t, it, prim, mesh, c = TD.run_case(ini)
    _, otext = TD.csv_rows(mesh, prim)
    assert text == otext
