"""The drop-in's output formats and restart on the GPU (`ccfd case.ini`,
`ccfd case.ini file.cgns`; src/main.c:53-94, src/output.c:333-525,
src/initialCondition.c:94-158).

* CGNS (outputFormat 1): the grid file, one flow solution per output and
  the master file listing them; a stationary run restarted from an
  intermediate solution reaches the uninterrupted run's final state and
  appends to the analysis file. The state is stored in full precision; the
  time is read back from the descriptor's 12 decimals (the reference's
  "%20.12f"), and the exact-function boundary states of this case depend
  on it, so the runs agree to about 1e-12 rather than bit for bit.
* Curve (outputFormat 2): the CSV's sorted rows as density, velocity and
  pressure blocks.
* A restart file of another mesh is refused.
* CGNS mesh input (meshFormat .cgns, readCGNS, src/mesh.c:1153-1394): the
  grid file of a run, read back as the mesh of the same case, gives the
  same solution bit for bit."""
import ctypes as C
import glob
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from ccfd_amd import _native as N

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIBDIR = os.path.join(ROOT, "ccfd_amd", "lib")
CCFD = os.path.join(LIBDIR, "ccfd")


def _cgns():
    path = os.path.join(LIBDIR, "libccfd_cgns.so")
    if not os.path.exists(path):
        pytest.skip("libccfd_cgns.so not built (no HDF5)")
    L = C.CDLL(path)
    L.ccfd_cgns_read_solution.argtypes = [
        C.c_char_p, C.c_long, N.c_double_p, C.POINTER(C.c_double),
        C.POINTER(C.c_double), C.c_char_p]
    return L


def _read(L, path, n):
    prim = np.zeros((n, 4))
    t, to = C.c_double(), C.c_double()
    err = C.create_string_buffer(256)
    rc = L.ccfd_cgns_read_solution(path.encode(), n, N.dptr(prim), C.byref(t),
                                   C.byref(to), err)
    assert rc == 0, err.value
    return prim, t.value


def _write(path, items):
    with open(path, "w") as f:
        for k, v in items:
            f.write("%s = %s\n" % (k, v))


def _run(d, *args, ok=True):
    print("run:", " ".join(args), "in", d, flush=True)
    r = subprocess.run([CCFD] + list(args), cwd=d, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=120)
    if ok:
        assert r.returncode == 0, r.stdout[-3000:]
    return r


def _stationary(d, nx=20, ny=12, fmt=1):
    ini = os.path.join(d, "st.ini")
    _write(ini, [
        ("meshType", 1), ("nElemsX", nx), ("nElemsY", ny),
        ("X0", "(/0.0, 0.0/)"), ("Xmax", "(/2.0, 1.2/)"),
        ("nBCsegments", "(/1, 1, 1, 1/)"),
        ("meshBCtype", 601), ("meshBCtype", 601), ("meshBCtype", 601),
        ("meshBCtype", 601),
        ("gamma", 1.4), ("fluxFunction", 5), ("CFL", 0.8), ("timeOrder", 2),
        ("nRKstages", 3), ("spatialOrder", 2), ("limiter", 2),
        ("venk_k", 10.0), ("stationary", "T"), ("abortResidual", "1e-30"),
        ("icType", 2), ("exactFunc", 3),
        ("nBC", 1), ("BCtype", 601), ("BCexactFunc", 3),
        ("fileName", "st"), ("IOtimeInterval", "1e9"),
        ("IOiterInterval", 20), ("outputFormat", fmt),
        ("tEnd", "1e9"), ("maxIter", 60)])
    return ini


def test_cgns_output_and_stationary_restart():
    L = _cgns()
    n = 20 * 12
    with tempfile.TemporaryDirectory() as d, tempfile.TemporaryDirectory() as d2:
        ini = _stationary(d)
        _run(d, ini)
        names = sorted(os.path.basename(p) for p in glob.glob(os.path.join(d, "st_*")))
        print(names)
        for want in ("st_mesh.cgns", "st_Master.cgns", "st_analysis.csv",
                     "st_000000000.cgns", "st_000000020.cgns",
                     "st_000000040.cgns", "st_000000060.cgns"):
            assert want in names
        full, t_full = _read(L, os.path.join(d, "st_000000060.cgns"), n)
        mid, t_mid = _read(L, os.path.join(d, "st_000000020.cgns"), n)
        assert t_mid > 0 and not np.array_equal(full, mid)
        # restart from iteration 20: the grid file and the first 20 lines
        # of the analysis file are there already
        shutil.copy(ini, d2)
        shutil.copy(os.path.join(d, "st_mesh.cgns"), d2)
        shutil.copy(os.path.join(d, "st_000000020.cgns"), d2)
        with open(os.path.join(d, "st_analysis.csv")) as f:
            lines = f.read().splitlines()
        with open(os.path.join(d2, "st_analysis.csv"), "w") as f:
            f.write("\n".join(lines[:21]) + "\n")
        mtime = os.path.getmtime(os.path.join(d2, "st_mesh.cgns"))
        out = _run(d2, "st.ini", "st_000000020.cgns").stdout
        print(out[-600:])
        # no grid file is written on a restart
        assert os.path.getmtime(os.path.join(d2, "st_mesh.cgns")) == mtime
        assert not os.path.exists(os.path.join(d2, "st_000000000.cgns"))
        again, t_again = _read(L, os.path.join(d2, "st_000000060.cgns"), n)
        print("restart vs uninterrupted: max |diff| %.3e, t %.15g %.15g" % (
            np.abs(again - full).max(), t_again, t_full))
        np.testing.assert_allclose(again, full, rtol=1e-10, atol=0)
        assert not np.array_equal(again, mid)
        assert abs(t_again - t_full) <= 1e-11 * abs(t_full)
        with open(os.path.join(d2, "st_analysis.csv")) as f:
            got = f.read().splitlines()
        assert len(got) == len(lines)
        assert got[:21] == lines[:21]       # appended to, not rewritten
        for a, b in zip(got[21:], lines[21:]):
            fa, fb = a.split(","), b.split(",")
            assert fa[0] == fb[0]
            np.testing.assert_allclose([float(x) for x in fa[1:]],
                                       [float(x) for x in fb[1:]], rtol=1e-8)


def test_restart_file_of_another_mesh_is_refused():
    _cgns()
    with tempfile.TemporaryDirectory() as d, tempfile.TemporaryDirectory() as d2:
        _run(d, _stationary(d))
        ini2 = _stationary(d2, nx=21)
        shutil.copy(os.path.join(d, "st_000000020.cgns"), d2)
        r = _run(d2, ini2, "st_000000020.cgns", ok=False)
        assert r.returncode != 0
        assert "Wrong Number of Elements" in r.stdout
        r = _run(d2, ini2, "st_999.cgns", ok=False)
        assert r.returncode != 0 and "No restart file found" in r.stdout


def test_curve_output_has_the_csv_rows():
    with tempfile.TemporaryDirectory() as d, tempfile.TemporaryDirectory() as d2:
        _run(d, _stationary(d, fmt=3))
        _run(d2, _stationary(d2, fmt=2))
        with open(os.path.join(d, "st_000000060.csv")) as f:
            csv = [r.split(",") for r in f.read().splitlines()[1:]]
        with open(os.path.join(d2, "st_000000060.curve")) as f:
            blocks = f.read().split("\n\n")
        assert [b.splitlines()[0] for b in blocks] == [
            "#Density", "#Velocity", "#Pressure"]
        assert len(csv) == 20 * 12
        for b, col in zip(blocks, (1, 2, 3)):
            rows = [r.split(",") for r in b.splitlines()[1:] if r]
            assert [r[0] for r in rows] == [c[0] for c in csv]
            assert [r[1] for r in rows] == [c[col] for c in csv]


def test_cgns_grid_file_as_mesh_input_reproduces_the_run():
    L = _cgns()
    n = 20 * 12
    mesh_keys = ("meshType", "nElemsX", "nElemsY", "X0", "Xmax",
                 "nBCsegments", "meshBCtype", "fileName")
    with tempfile.TemporaryDirectory() as d:
        _run(d, _stationary(d))
        full, t_full = _read(L, os.path.join(d, "st_000000060.cgns"), n)
        with open(os.path.join(d, "st.ini")) as f:
            lines = [x for x in f.read().splitlines()
                     if x.split("=")[0].strip() not in mesh_keys]
        lines += ["meshType = 2", "meshFormat = .cgns", "meshFile = st_mesh",
                  "fileName = rd"]
        with open(os.path.join(d, "rd.ini"), "w") as f:
            f.write("\n".join(lines) + "\n")
        _run(d, "rd.ini")
        got, t_got = _read(L, os.path.join(d, "rd_000000060.cgns"), n)
        assert t_got == t_full
        assert np.array_equal(got, full)
