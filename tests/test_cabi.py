"""The C ABI libraries load and export every function include/*.h declares
(no device calls: runs without a GPU)."""
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
    ("ccfd_mesh.h", "libccfd_host.so"),
    ("ccfd_dropin.h", "libccfd_host.so"),
])
def test_library_exports_declared_symbols(header, lib):
    if not os.path.exists(os.path.join(ROOT, "include", header)):
        pytest.skip(header + " not present")
    path = os.path.join(LIB, lib)
    assert os.path.exists(path), path + " not built"
    L = C.CDLL(path)
    names = declared(header)
    assert len(names) > 3
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_every_header_is_checked():
    hdrs = sorted(f for f in os.listdir(os.path.join(ROOT, "include"))
                  if f.endswith(".h"))
    assert hdrs == ["ccfd_batch.h", "ccfd_dropin.h", "ccfd_gpu.h",
                    "ccfd_mesh.h"]
