"""The drop-in driver's set-up from a parameter file (ccfd_init,
ccfd_amd/csrc/driver.c; the reference's init chain, src/main.c:53-121), up
to the point where it needs a device: every committed case file reads, and
its mesh, boundary conditions and initial state build; removing any key
fails with "missing ini key <key>" exactly when the reference reads that key
without a default (a NULL proposal to readInTools' getters, found in the
reference's sources), and otherwise falls back to the default. Host only."""
import ctypes as C
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "ccfd_amd", "lib", "libccfd_host.so")
FILES = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "**", "*.ini"),
                         recursive=True))
REF_SRC = "/root/reference/src"


def _lib():
    if not os.path.exists(HOST):
        pytest.skip("libccfd_host.so not built")
    L = C.CDLL(HOST)
    L.ccfd_init.argtypes = [C.c_char_p, C.c_int, C.c_char_p]
    L.ccfd_init.restype = C.c_int
    return L


def _init(L, path):
    """ccfd_init's message; an initialised case (a GPU is present) is
    finalised again"""
    err = C.create_string_buffer(256)
    rc = L.ccfd_init(path.encode(), 0, err)
    if rc == 0:
        L.ccfd_finalize()
        return ""
    return err.value.decode()


def _device_only(msg):
    """set-up got as far as the device"""
    return msg == "" or "device" in msg.lower() or "hip" in msg.lower()


def _key(line):
    head = re.split(r"[!#=]", line, maxsplit=1)[0]
    return "".join(ch for ch in head if ch.isalnum() or ch == "_").lower()


def _reference_mandatory():
    src = "".join(open(f).read() for f in glob.glob(os.path.join(REF_SRC, "*.c")))
    pat = (r'get(?:Int|Dbl|Str|Bool|IntArray|DblArray)\(\s*"(\w+)"\s*,'
           r'\s*(?:[^,()"]+,\s*)?NULL\s*\)')
    return {k.lower() for k in re.findall(pat, src)}


def test_every_case_file_sets_up(tmp_path, monkeypatch):
    L = _lib()
    monkeypatch.chdir(tmp_path)
    assert len(FILES) >= 15
    for path in FILES:
        msg = _init(L, path)
        assert _device_only(msg), (os.path.basename(path), msg)


def test_missing_keys_fail_as_in_the_reference(tmp_path, monkeypatch):
    L = _lib()
    if not os.path.isdir(REF_SRC):
        pytest.skip("reference sources not present")
    mand = _reference_mandatory()
    assert {"fluxfunction", "nbc", "tend", "nbcsegments"} <= mand
    monkeypatch.chdir(tmp_path)
    checked = 0
    for path in FILES:
        lines = open(path).read().splitlines()
        keys = list(dict.fromkeys(k for k in map(_key, lines) if k))
        for k in keys:
            cut = tmp_path / "cut.ini"
            cut.write_text("\n".join(x for x in lines if _key(x) != k) + "\n")
            msg = _init(L, str(cut))
            name = os.path.basename(path)
            if k in mand:
                assert msg.lower() == "missing ini key " + k, (name, k, msg)
            else:
                assert _device_only(msg), (name, k, msg)
            checked += 1
    assert checked > 200


def test_command_line_errors(tmp_path):
    """`ccfd case.ini` (src/main.c:35-130): wrong usage, an unreadable file
    and a missing mandatory key end with status 1 and a message, before any
    device is touched"""
    import subprocess
    exe = os.path.join(ROOT, "ccfd_amd", "lib", "ccfd")
    if not os.path.exists(exe):
        pytest.skip("ccfd not built")

    def run(*args):
        return subprocess.run([exe, *args], capture_output=True, text=True,
                              timeout=60, cwd=tmp_path)
    r = run()
    assert r.returncode == 1 and "usage" in r.stdout
    r = run(str(tmp_path / "absent.ini"))
    assert r.returncode == 1 and "could not open parameter file" in r.stdout
    src = [x for x in open(FILES[0]).read().splitlines()
           if _key(x) != "fluxfunction"]
    (tmp_path / "cut.ini").write_text("\n".join(src) + "\n")
    r = run("cut.ini")
    assert r.returncode == 1
    assert "missing ini key fluxFunction" in r.stdout + r.stderr


def test_unsupported_settings_are_refused(tmp_path, monkeypatch):
    """settings outside this path, or incomplete, end set-up with a
    message, not a wrong run: a source term without its function (the
    reference's mandatory sourceFunction, src/equation.c:121), an
    exact-function initial condition the reference does not know, a gmsh
    file of another format, an unknown mesh format and a missing CGNS mesh
    file"""
    L = _lib()
    monkeypatch.chdir(tmp_path)
    base = open(FILES[0]).read()
    (tmp_path / "src.ini").write_text(base + "\ncalcSource = T\n")
    assert _init(L, str(tmp_path / "src.ini")) == \
        "missing ini key sourceFunction"
    lines = [x for x in base.splitlines() if _key(x) != "exactfunc"]
    (tmp_path / "ef.ini").write_text("\n".join(lines + ["exactFunc = 7"]) + "\n")
    assert _init(L, str(tmp_path / "ef.ini")) == "unknown exact function 7"
    (tmp_path / "m3.msh").write_text("$MeshFormat\n3.0 0 8\n$EndMeshFormat\n")
    lines = [x for x in base.splitlines()
             if _key(x) not in ("meshtype", "meshformat", "meshfile")]
    (tmp_path / "m3.ini").write_text("\n".join(
        lines + ["meshType = 2", "meshFormat = .msh", "meshFile = m3"]) + "\n")
    msg = _init(L, str(tmp_path / "m3.ini"))
    assert "format 3 is not supported" in msg, msg
    (tmp_path / "xy.ini").write_text("\n".join(
        lines + ["meshType = 2", "meshFormat = .xyz", "meshFile = m3"]) + "\n")
    msg = _init(L, str(tmp_path / "xy.ini"))
    assert "Unknown Mesh Format" in msg, msg
    (tmp_path / "cg.ini").write_text("\n".join(
        lines + ["meshType = 2", "meshFormat = .cgns", "meshFile = m3"]) + "\n")
    msg = _init(L, str(tmp_path / "cg.ini"))
    assert "cannot open" in msg or "libccfd_cgns.so" in msg, msg


def test_cgns_grid_file_is_read_as_a_mesh(tmp_path, monkeypatch):
    """meshFormat = .cgns (readCGNS, src/mesh.c:1153-1394): set-up writes
    the grid file of a CGNS-output case before it reaches the device, and a
    case naming that file as its mesh reads it and gets as far as the
    device"""
    from tests.test_gpu_cgns_restart import _stationary
    L = _lib()
    if not os.path.exists(os.path.join(os.path.dirname(HOST), "libccfd_cgns.so")):
        pytest.skip("libccfd_cgns.so not built")
    monkeypatch.chdir(tmp_path)
    ini = _stationary(str(tmp_path))
    assert _device_only(_init(L, ini))
    assert (tmp_path / "st_mesh.cgns").exists()
    keys = ("meshtype", "nelemsx", "nelemsy", "x0", "xmax", "nbcsegments",
            "meshbctype", "filename")
    lines = [x for x in open(ini).read().splitlines() if _key(x) not in keys]
    (tmp_path / "rd.ini").write_text("\n".join(lines + [
        "meshType = 2", "meshFormat = .cgns", "meshFile = st_mesh",
        "fileName = rd"]) + "\n")
    msg = _init(L, str(tmp_path / "rd.ini"))
    assert _device_only(msg), msg
    (tmp_path / "st_mesh.cgns").write_bytes(b"not hdf5")
    msg = _init(L, str(tmp_path / "rd.ini"))
    assert "cannot open" in msg, msg
