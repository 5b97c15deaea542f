"""The drop-in host's parameter-file reader (ccfd_amd/csrc/ini.c) against
the reference's own readInTools (src/readInTools.c:38-405, compiled into
oracle/_ref/libccfdref.so): every committed case file and a file of awkward
lines are read key by key, in file order, with both. The values (string,
double, integer, bool, arrays), the positional reading of repeated keys,
the key counts and the defaults of absent keys must agree. Host only; skips
where the reference harness is not built."""
import ctypes as C
import glob
import os
import re

import pytest

from oracle import bindings as B

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "ccfd_amd", "lib", "libccfd_host.so")
FILES = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "**", "*.ini"),
                         recursive=True))

AWKWARD = """\
  Mesh File = (/ mesh_01.msh /)   ! trailing comment
CFL=0.9#no space before the comment
fluxFunction = 5
BCtype = 1
BCtype = 2
BCtype = 4
rho = 1.0
rho = 0.125
Stop_Time = 2.5e-1
iniFile   =   "quoted value"
IC_vector = (/1.0, -0.5, 3e-2/)
  !commented = 3
# also a comment
weird!key = 4
=novalue
key_only
Flag = T
Flag2 = true
Flag3 = .TRUE.
Flag4 = F
nIter = 1.9e3
Spaced Key Name = 7
"""


def _libs():
    if not B.ref_available():
        pytest.skip("reference harness not built")
    if not os.path.exists(HOST):
        pytest.skip("libccfd_host.so not built")
    ref = C.CDLL(os.path.join(B.HERE, "_ref", "libccfdref.so"))
    ref.fillCmds.argtypes = [C.c_char_p]
    ref.getStr.restype = C.c_char_p
    ref.getStr.argtypes = [C.c_char_p, C.c_char_p]
    ref.getDbl.restype = C.c_double
    ref.getDbl.argtypes = [C.c_char_p, C.c_char_p]
    ref.getInt.restype = C.c_long
    ref.getInt.argtypes = [C.c_char_p, C.c_char_p]
    ref.getBool.restype = C.c_bool
    ref.getBool.argtypes = [C.c_char_p, C.c_char_p]
    ref.getDblArray.restype = C.POINTER(C.c_double)
    ref.getDblArray.argtypes = [C.c_char_p, C.c_int, C.c_char_p]
    ref.countKeys.restype = C.c_int
    ref.countKeys.argtypes = [C.c_char_p, C.c_int]
    host = C.CDLL(HOST)
    host.ccfd_ini_read.restype = C.c_void_p
    host.ccfd_ini_read.argtypes = [C.c_char_p, C.c_char_p]
    host.ccfd_ini_free.argtypes = [C.c_void_p]
    for f in ("str", "dbl", "long", "bool", "dbl_array", "count"):
        getattr(host, "ccfd_ini_" + f).restype = C.c_int
    host.ccfd_ini_str.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p,
                                  C.c_char_p, C.c_size_t]
    host.ccfd_ini_dbl.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p,
                                  C.POINTER(C.c_double)]
    host.ccfd_ini_long.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p,
                                   C.POINTER(C.c_long)]
    host.ccfd_ini_bool.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p,
                                   C.POINTER(C.c_int)]
    host.ccfd_ini_dbl_array.argtypes = [C.c_void_p, C.c_char_p, C.c_int,
                                        C.c_char_p, C.POINTER(C.c_double)]
    host.ccfd_ini_count.argtypes = [C.c_void_p, C.c_char_p]
    return ref, host


def _keys(path):
    """the keys of a file in order (letters, digits and '_' before the first
    of '!', '#', '='), as both readers store them"""
    out = []
    for line in open(path):
        head = re.split(r"[!#=]", line, maxsplit=1)[0]
        k = "".join(ch for ch in head if ch.isalnum() or ch == "_")
        if k:
            out.append(k)
    return out


class _Pair:
    """one file opened by both readers"""

    def __init__(self, ref, host, path):
        self.ref, self.host = ref, host
        C.c_void_p.in_dll(ref, "firstCmd").value = None
        ref.fillCmds(path.encode())
        err = C.create_string_buffer(256)
        self.h = host.ccfd_ini_read(path.encode(), err)
        assert self.h, err.value

    def close(self):
        C.CDLL(None).fflush(None)    # the reference's printf, while captured
        self.ref.freeCmds()
        C.c_void_p.in_dll(self.ref, "firstCmd").value = None
        self.host.ccfd_ini_free(self.h)

    def count(self, key):
        k = key.encode()
        n = self.host.ccfd_ini_count(self.h, k)
        assert self.ref.countKeys(k, 0) == n, key
        return n

    def string(self, key, default=b"absent"):
        buf = C.create_string_buffer(256)
        assert self.host.ccfd_ini_str(self.h, key.encode(), default, buf, 256) == 0
        return self.ref.getStr(key.encode(), default), buf.value

    def number(self, key, default=b"-7.5"):
        d = C.c_double()
        assert self.host.ccfd_ini_dbl(self.h, key.encode(), default, C.byref(d)) == 0
        return self.ref.getDbl(key.encode(), default), d.value

    def integer(self, key, default=b"-3"):
        v = C.c_long()
        assert self.host.ccfd_ini_long(self.h, key.encode(), default, C.byref(v)) == 0
        return self.ref.getInt(key.encode(), default), v.value

    def flag(self, key, default=b"F"):
        v = C.c_int()
        assert self.host.ccfd_ini_bool(self.h, key.encode(), default, C.byref(v)) == 0
        return bool(self.ref.getBool(key.encode(), default)), bool(v.value)

    def array(self, key, n, default):
        out = (C.c_double * n)()
        assert self.host.ccfd_ini_dbl_array(self.h, key.encode(), n, default,
                                            out) == 0
        r = self.ref.getDblArray(key.encode(), n, default)
        return [r[i] for i in range(n)], list(out)


def _write_awkward(tmp_path):
    p = tmp_path / "awkward.ini"
    p.write_text(AWKWARD)
    return str(p)


@pytest.mark.parametrize("reader", ["string", "number", "integer", "flag"])
def test_every_case_file_reads_alike(reader, tmp_path):
    ref, host = _libs()
    files = FILES + [_write_awkward(tmp_path)]
    assert len(files) >= 15
    for path in files:
        pair = _Pair(ref, host, path)
        try:
            keys = _keys(path)
            # mixed case on purpose: keys are case-insensitive
            for i, key in enumerate(keys):
                k = key.upper() if i % 2 else key
                n = pair.count(k)
                assert n >= 1
                want, got = getattr(pair, reader)(k)
                assert want == got, (os.path.basename(path), key, want, got)
                assert pair.count(k) == n - 1        # consumed
            # every key is used up: the defaults come back
            for key in set(keys):
                want, got = getattr(pair, reader)(key)
                assert want == got, (key, want, got)
        finally:
            pair.close()


def test_awkward_lines(tmp_path):
    ref, host = _libs()
    pair = _Pair(ref, host, _write_awkward(tmp_path))
    try:
        # repeated keys are read in file order
        assert pair.integer("BCtype") == (1, 1)
        assert pair.integer("bctype") == (2, 2)
        assert pair.string("meshfile") == (b"mesh_01.msh", b"mesh_01.msh")
        assert pair.number("cfl") == (0.9, 0.9)
        assert pair.array("ic_vector", 3, b"0,0,0") == ([1.0, -0.5, 0.03],
                                                        [1.0, -0.5, 0.03])
        assert pair.string("inifile") == (b"quotedvalue", b"quotedvalue")
        assert pair.integer("niter") == (1900, 1900)
        assert pair.flag("flag") == (True, True)
        assert pair.flag("flag2") == (True, True)
        assert pair.flag("flag3") == (False, False)   # only T/t/True/true
        assert pair.string("weird") == (b"", b"")      # value cut at '!'
        assert pair.string("key_only") == (b"", b"")
        assert pair.integer("spacedkeyname") == (7, 7)
        assert pair.array("absent", 2, b"1.5,2.5") == ([1.5, 2.5], [1.5, 2.5])
    finally:
        pair.close()


def test_oracle_ini_reads_alike(tmp_path):
    """the oracle's Python restatement of the reader (oracle/timedisc.py,
    used by its make check runs) takes the same keys and values, in the
    same positional order, as the reference's readInTools"""
    from oracle import timedisc as TD
    ref, host = _libs()
    for path in FILES + [_write_awkward(tmp_path)]:
        pair = _Pair(ref, host, path)
        ini = TD.Ini(path)
        try:
            for key in _keys(path):
                want, _ = pair.string(key)
                assert ini.str(key).encode() == want, (path, key)
            assert ini.cmds == []
            assert ini.str("absent", "dflt") == "dflt"
        finally:
            pair.close()
