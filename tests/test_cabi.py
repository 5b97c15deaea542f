"""The C ABI libraries load and export every function include/*.h declares;
every entry point of the HIP library rejects a null handle or bad arguments
with CCFD_EINVAL and a message before it touches a device; and without a
device the product path raises instead of computing on the CPU. Runs
without a GPU (no compute calls; the last test skips where a device is
present)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "ccfd_amd", "lib")


def declared(header):
    src = open(os.path.join(ROOT, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"^\w[\w\s\*]*?\b(\w+)\s*\(", src, re.M)))


@pytest.mark.parametrize("header,lib", [
    ("ccfd_gpu.h", "libccfd_gpu.so"),
    ("ccfd_batch.h", "libccfd_gpu.so"),
    ("ccfd_serve.h", "libccfd_gpu.so"),
    ("ccfd_mesh.h", "libccfd_host.so"),
    ("ccfd_dropin.h", "libccfd_host.so"),
    ("ccfd_cgns.h", "libccfd_cgns.so"),
])
def test_library_exports_declared_symbols(header, lib):
    if not os.path.exists(os.path.join(ROOT, "include", header)):
        pytest.skip(header + " not present")
    path = os.path.join(LIB, lib)
    if lib == "libccfd_cgns.so" and not os.path.exists(path):
        pytest.skip("built only where HDF5 is installed")
    assert os.path.exists(path), path + " not built"
    L = C.CDLL(path)
    names = declared(header)
    assert len(names) > 3
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_every_header_is_checked():
    hdrs = sorted(f for f in os.listdir(os.path.join(ROOT, "include"))
                  if f.endswith(".h"))
    assert hdrs == ["ccfd_batch.h", "ccfd_cgns.h", "ccfd_dropin.h",
                    "ccfd_gpu.h", "ccfd_mesh.h", "ccfd_serve.h"]


def _entry_points(header):
    """(name, [parameter declarations]) of every int-returning function"""
    src = open(os.path.join(ROOT, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    out = []
    for name, args in re.findall(r"^int\s+(\w+)\s*\(([^)]*)\)", src, re.M):
        args = [a.strip() for a in args.split(",")]
        out.append((name, [a for a in args if a and a != "void"]))
    return out


def _null_args(params):
    """NULL for every pointer, 1 for every integer, 0.0 for every double"""
    vals = []
    for a in params:
        if "*" in a or "[" in a:
            vals.append(C.c_void_p(None))
        elif "double" in a:
            vals.append(C.c_double(0.0))
        elif "size_t" in a or "long" in a:
            vals.append(C.c_long(1))
        else:
            vals.append(C.c_int(1))
    return vals


@pytest.mark.parametrize("header", ["ccfd_gpu.h", "ccfd_batch.h",
                                    "ccfd_serve.h"])
def test_null_handles_are_rejected(header):
    L = C.CDLL(os.path.join(LIB, "libccfd_gpu.so"))
    L.ccfd_gpu_last_error.restype = C.c_char_p
    checked = 0
    for name, params in _entry_points(header):
        first = params[0] if params else ""
        if not ("ccfd_gpu_ctx" in first or "ccfd_batch" in first or
                "ccfd_server" in first):
            continue
        f = getattr(L, name)
        f.restype = C.c_int
        rc = f(*_null_args(params))
        if name.endswith("_destroy"):
            assert rc == 0, name      # like free(NULL)
        else:
            assert rc == -1, (name, rc)  # CCFD_EINVAL
            assert L.ccfd_gpu_last_error(), name
        checked += 1
    assert checked >= (5 if header == "ccfd_serve.h" else 15)


def test_context_free_entry_points_check_arguments_first():
    L = C.CDLL(os.path.join(LIB, "libccfd_gpu.so"))
    L.ccfd_gpu_last_error.restype = C.c_char_p
    L.ccfd_gpu_riemann_batch.argtypes = [C.c_int, C.c_int, C.c_double, C.c_long,
                                         C.c_void_p, C.c_void_p, C.c_void_p]
    assert L.ccfd_gpu_riemann_batch(12, 0, 1.4, 4, None, None, None) == -1
    assert b"unknown flux" in L.ccfd_gpu_last_error()
    assert L.ccfd_gpu_riemann_batch(5, 0, 1.4, 4, None, None, None) == -1
    assert L.ccfd_gpu_riemann_batch(5, 0, 1.4, -1, None, None, None) == -1
    assert L.ccfd_gpu_riemann_batch(5, 0, 1.4, 0, None, None, None) == 0
    L.ccfd_gpu_copy_probe.argtypes = [C.c_int, C.c_size_t, C.c_int, C.c_void_p]
    assert L.ccfd_gpu_copy_probe(0, 1024, 1, None) == -1


def test_no_cpu_fallback_without_a_device():
    """with no HIP device the product path raises the library's message: a
    context, a batch and a flux batch all fail, none computes on the CPU
    (skipped where a device is present)"""
    import numpy as np
    from ccfd_amd import batch as BT
    from ccfd_amd import gpu as G
    from tests import cases
    L = C.CDLL(os.path.join(LIB, "libccfd_gpu.so"))
    h = C.c_void_p()
    if L.ccfd_gpu_create(C.byref(h), 0) == 0:
        L.ccfd_gpu_destroy(h)
        pytest.skip("a HIP device is present")
    mesh = cases.sod_mesh(20)
    with pytest.raises(G.GpuError, match="device"):
        G.Solver(mesh, G.params())
    with pytest.raises(Exception, match="device"):
        BT.Batch(G.params(), max_sides=4)
    state = np.tile([1.0, 0.1, 0.0, 1.0], (4, 1))
    with pytest.raises(G.GpuError):
        G.riemann_batch(5, state, state)
