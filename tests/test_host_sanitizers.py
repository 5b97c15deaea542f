"""The host C library (ccfd_amd/csrc/mesh.c, ini.c, driver.c) built with
AddressSanitizer, UndefinedBehaviorSanitizer and leak detection (gcc, host
code only), driven by tests/native/host_sanitize.c: every committed case
file through ccfd_init (which stops where it would create the device
context, freeing what it built), a missing file, and the gmsh meshes of the
reference's own cases read, built, refused when opened up, and freed; and
the host C++ of the device library (tests/native/host_sanitize_cxx.cpp):
device layout, halo lists and multigrid levels of those meshes, and the
batched arena's allocator through random sequences; and the CGNS library
(tests/native/cgns_sanitize.c). Any sanitizer report fails the test. Host
only."""
import glob
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "ccfd_amd", "csrc")
LIB = os.path.join(ROOT, "ccfd_amd", "lib")


SAN = ["-g", "-O1", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
       "-fno-sanitize-recover=undefined"]
LINK = ["-L" + LIB, "-lccfd_gpu", "-Wl,-rpath," + LIB,
        "-Wl,-rpath-link,/opt/rocm/lib", "-lm"]
# the environment is kept as it is; ASan is told not to insist on being the
# first library loaded
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1:"
                                    "verify_asan_link_order=0",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
MESHES = sorted(glob.glob("/root/reference/calc/**/*.msh", recursive=True) +
                glob.glob("/root/reference/calc/**/*.mesh", recursive=True))


def _host_library(tmp_path):
    """libccfd_host.so with the sanitizers, in tmp_path"""
    cc = shutil.which("gcc")
    if not cc or not os.path.exists(os.path.join(LIB, "libccfd_gpu.so")):
        pytest.skip("gcc or libccfd_gpu.so missing")
    srcs = [os.path.join(CSRC, s) for s in ("mesh.c", "ini.c", "driver.c")]
    r = subprocess.run([cc, "-std=c99", "-D_POSIX_C_SOURCE=200809L",
                        "-ffp-contract=off", "-fPIC", "-shared",
                        "-I" + os.path.join(ROOT, "include"), "-o",
                        str(tmp_path / "libccfd_host.so")]
                       + SAN + srcs + LINK, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    return cc


def _clean(r):
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-3000:]
    assert "Sanitizer" not in out and "runtime error" not in out, out[-3000:]
    return out


def test_host_library_under_sanitizers(tmp_path):
    cc = _host_library(tmp_path)
    exe = tmp_path / "host_sanitize"
    r = subprocess.run([cc, "-std=c99", "-I" + os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "native", "host_sanitize.c"),
                        "-L" + str(tmp_path), "-lccfd_host",
                        "-Wl,-rpath," + str(tmp_path), "-o", str(exe)]
                       + SAN + LINK, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    inis = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "*", "*.ini")))
    r = subprocess.run([str(exe)] + inis + ["--"] + MESHES,
                       capture_output=True, text=True, timeout=600,
                       env=ENV, cwd=tmp_path)
    out = _clean(r)
    assert out.count("init ") == len(inis) and "ACCEPTED" not in out
    assert "missing file: -1" in out


def test_device_library_host_code_under_sanitizers(tmp_path):
    _host_library(tmp_path)
    cxx = shutil.which("g++")
    if not cxx:
        pytest.skip("g++ missing")
    exe = tmp_path / "host_sanitize_cxx"
    r = subprocess.run([cxx, "-std=c++17", "-I" + os.path.join(ROOT, "include"),
                        "-I" + CSRC,
                        os.path.join(ROOT, "tests", "native",
                                     "host_sanitize_cxx.cpp"),
                        os.path.join(CSRC, "soa.cpp"),
                        os.path.join(CSRC, "agglomerate.cpp"),
                        "-L" + str(tmp_path), "-lccfd_host",
                        "-Wl,-rpath," + str(tmp_path), "-o", str(exe)]
                       + SAN + LINK, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)] + MESHES, capture_output=True, text=True,
                       timeout=600, env=ENV, cwd=tmp_path)
    out = _clean(r)
    assert out.count("allocator ") == 3
    assert "used 0 high 0 ranges 1" in out
    # every mesh file of the reference's cases (gmsh 2 and 4, EMC2) reads
    assert out.count(" levels") == len(MESHES)


def test_cgns_library_under_sanitizers(tmp_path):
    """libccfd_cgns.so (ccfd_amd/csrc/cgns_hdf5.c) with the sanitizers,
    driven by tests/native/cgns_sanitize.c: every entry point and its
    refusals on a small mixed mesh"""
    import sys
    sys.path.insert(0, ROOT)
    from ccfd_amd import build as BLD
    pre = BLD.hdf5_prefix()
    if pre is None:
        pytest.skip("HDF5 not available")
    cc = _host_library(tmp_path)
    # HDF5 by its path: a -L to its prefix would also offer that prefix's
    # own sanitizer runtimes to the linker
    h5 = ["-I" + os.path.join(pre, "include"),
          os.path.join(pre, "lib", "libhdf5.so"),
          "-Wl,-rpath," + os.path.join(pre, "lib")]
    r = subprocess.run([cc, "-std=c99", "-D_POSIX_C_SOURCE=200809L", "-fPIC",
                        "-shared", "-I" + os.path.join(ROOT, "include"), "-o",
                        str(tmp_path / "libccfd_cgns.so"),
                        os.path.join(CSRC, "cgns_hdf5.c")] + SAN + h5,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    exe = tmp_path / "cgns_sanitize"
    r = subprocess.run([cc, "-std=c99", "-I" + os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "native", "cgns_sanitize.c"),
                        "-L" + str(tmp_path), "-lccfd_cgns", "-lccfd_host",
                        "-Wl,-rpath," + str(tmp_path), "-o", str(exe)]
                       + SAN + LINK, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True,
                       timeout=300, env=ENV, cwd=tmp_path)
    out = _clean(r)
    print(out)
    for line in ("write_mesh: 0", "write_solution: 0", "write_master: 0",
                 "read_solution: 0 same 1 t 1",
                 "read_mesh g.cgns: 6 nodes 2 tri 1 quad 6 edges",
                 "read_mesh s_1.cgns: 6 nodes 2 tri 1 quad 0 edges",
                 "null: -1"):
        assert line in out, line
    assert "Wrong Number of Elements" in out
    assert "read_mesh absent.cgns: -1 cannot open" in out
