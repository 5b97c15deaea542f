"""ctypes bindings of the native libraries.

The product path is libccfd_host.so (C host code) and libccfd_gpu.so (HIP
kernels behind the C ABI in include/ccfd_gpu.h). Both are loaded from
ccfd_amd/lib; a missing library raises instead of falling back to anything.
"""
import ctypes as C
import os

import numpy as np

LIBDIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")

c_double_p = C.POINTER(C.c_double)
c_long_p = C.POINTER(C.c_long)
c_int_p = C.POINTER(C.c_int)
c_uchar_p = C.POINTER(C.c_ubyte)


class CcfdBC(C.Structure):
    """struct ccfd_bc (include/ccfd_mesh.h)"""
    _fields_ = [("type", C.c_int), ("id", C.c_int), ("exactFunc", C.c_int),
                ("pVar", C.c_double * 4), ("connection", C.c_double * 2)]


class CcfdRawMesh(C.Structure):
    """struct ccfd_raw_mesh (include/ccfd_mesh.h)"""
    _fields_ = [
        ("nVert", C.c_long), ("xy", c_double_p),
        ("nTri", C.c_long), ("tri", c_long_p),
        ("nQuad", C.c_long), ("quad", c_long_p),
        ("nBCedge", C.c_long), ("bcEdge", c_long_p),
        ("nBC", C.c_int), ("bc", C.POINTER(CcfdBC)),
    ]


class CcfdMesh(C.Structure):
    """struct ccfd_mesh (include/ccfd_mesh.h)"""
    _fields_ = [
        ("nNodes", C.c_long), ("xy", c_double_p),
        ("nElems", C.c_long), ("nTrias", C.c_long), ("nQuads", C.c_long),
        ("elemType", c_int_p), ("domain", c_int_p), ("elemNode", c_long_p),
        ("bary", c_double_p), ("area", c_double_p), ("areaq", c_double_p),
        ("sx", c_double_p), ("sy", c_double_p), ("totalArea_q", C.c_double),
        ("nSlots", C.c_long), ("slotStart", c_long_p), ("nbr", c_long_p),
        ("partner", c_long_p), ("owner", c_uchar_p),
        ("n", c_double_p), ("len", c_double_p), ("fn", c_double_p),
        ("flen", c_double_p), ("GP", c_double_p), ("GPc", c_double_p),
        ("w", c_double_p), ("bbv", c_double_p), ("bbd", c_double_p),
        ("slotNode", c_long_p),
        ("nFaces", C.c_long), ("faceSlot", c_long_p),
        ("nGhosts", C.c_long), ("ghostSlot", c_long_p), ("ghostBC", c_int_p),
        ("ghostBary", c_double_p), ("ghostGP", c_double_p),
        ("nBC", C.c_int), ("bc", C.POINTER(CcfdBC)),
        ("xMin", C.c_double), ("xMax", C.c_double),
        ("yMin", C.c_double), ("yMax", C.c_double),
    ]


_libs = {}


def _load(name):
    if name in _libs:
        return _libs[name]
    path = os.path.join(LIBDIR, name)
    if not os.path.exists(path):
        raise RuntimeError(
            "%s is not built (run python -m ccfd_amd.build or "
            "__graft_entry__.build())" % path)
    lib = C.CDLL(path, mode=C.RTLD_GLOBAL)
    _libs[name] = lib
    return lib


def host():
    lib = _load("libccfd_host.so")
    if not getattr(lib, "_typed", False):
        lib.ccfd_mesh_build.restype = C.POINTER(CcfdMesh)
        lib.ccfd_mesh_build.argtypes = [
            C.c_long, c_double_p, C.c_long, c_long_p, C.c_long, c_long_p,
            C.c_long, c_long_p, C.c_int, C.POINTER(CcfdBC), C.c_char_p]
        lib.ccfd_mesh_free.argtypes = [C.POINTER(CcfdMesh)]
        lib.ccfd_mesh_free.restype = None
        lib.ccfd_cart_connectivity.restype = C.c_int
        lib.ccfd_cart_connectivity.argtypes = [
            C.c_long, C.c_long, C.c_double, C.c_double, C.c_double,
            C.c_double, c_int_p, c_int_p, c_int_p,
            c_long_p, C.POINTER(c_double_p), c_long_p, C.POINTER(c_long_p),
            c_long_p, C.POINTER(c_long_p)]
        lib.ccfd_gmsh_write.restype = C.c_int
        lib.ccfd_gmsh_write.argtypes = [
            C.c_char_p, C.c_long, c_double_p, C.c_long, c_long_p, C.c_long,
            c_long_p, C.c_long, c_long_p]
        for name in ("ccfd_gmsh_read", "ccfd_emc2_read"):
            f = getattr(lib, name)
            f.restype = C.c_int
            f.argtypes = [
                C.c_char_p, c_long_p, C.POINTER(c_double_p), c_long_p,
                C.POINTER(c_long_p), c_long_p, C.POINTER(c_long_p), c_long_p,
                C.POINTER(c_long_p), C.c_char_p]
        lib._typed = True
    return lib


libc = C.CDLL(None)
libc.free.argtypes = [C.c_void_p]
libc.free.restype = None


def dptr(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(c_double_p)


def lptr(a):
    assert a.dtype == np.int64 and a.flags.c_contiguous
    return a.ctypes.data_as(c_long_p)


def iptr(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(c_int_p)


def view(ptr, n, dtype):
    """copy n items behind a ctypes pointer into a numpy array"""
    if n == 0 or not ptr:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)
