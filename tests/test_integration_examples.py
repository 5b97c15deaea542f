"""The C examples in INTEGRATION.md compile against include/*.h: the
complete program links against the two libraries (it is not run: it needs a
GPU), and the fragments compile inside a function that declares what they
leave open. A renamed field, a changed signature or a missing include shows
up here. Host only."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DOC = os.path.join(ROOT, "INTEGRATION.md")
LIB = os.path.join(ROOT, "ccfd_amd", "lib")
INC = os.path.join(ROOT, "include")
CFLAGS = ["-std=c99", "-Wall", "-Werror", "-I" + INC]

# what the fragments take from their surroundings
PROLOGUE = """\
void example(long nVert, const double *xy, long nTri, const long *tri,
             long nQuad, const long *quad, long nBCedge, const long *bcEdge,
             int nBC, const ccfd_bc *bcs, double *prim, double *ut,
             double pTime, double tEnd)
{
	char err[256];
"""


def _c_blocks():
    with open(DOC) as f:
        return re.findall(r"```c\n(.*?)```", f.read(), re.S)


def _gcc():
    cc = shutil.which("gcc")
    if not cc:
        pytest.skip("gcc not available")
    return cc


def test_the_program_links(tmp_path):
    cc = _gcc()
    for name in ("libccfd_host.so", "libccfd_gpu.so"):
        if not os.path.exists(os.path.join(LIB, name)):
            pytest.skip(name + " not built")
    progs = [b for b in _c_blocks() if "int main" in b]
    assert len(progs) == 1
    src = tmp_path / "app.c"
    src.write_text(progs[0])
    r = subprocess.run(
        [cc] + CFLAGS + [str(src), "-L" + LIB, "-lccfd_host", "-lccfd_gpu",
                         "-Wl,-rpath," + LIB, "-Wl,-rpath-link,/opt/rocm/lib",
                         "-lm", "-o", str(tmp_path / "app")],
        capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr


def test_the_fragments_compile(tmp_path):
    cc = _gcc()
    frags = [b for b in _c_blocks() if "int main" not in b]
    assert len(frags) >= 3
    # the first fragment sets up what the others use (ctx, p, m); each of
    # the others gets a block of its own, as it would in a program
    includes, bodies = [], []
    for k, b in enumerate(frags):
        lines = b.splitlines()
        includes += [x for x in lines if x.startswith("#include")]
        body = [x for x in lines if not x.startswith("#include")]
        bodies += body if k == 0 else ["{"] + body + ["}"]
    src = tmp_path / "fragments.c"
    src.write_text("#include <stdio.h>\n" + "\n".join(dict.fromkeys(includes))
                   + "\n" + PROLOGUE + "\n".join(bodies) + "\n}\n")
    r = subprocess.run([cc] + CFLAGS + ["-c", str(src), "-o",
                                        str(tmp_path / "fragments.o")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr


def test_the_python_examples_use_the_real_api():
    """every ccfd_amd attribute the Python examples name exists, and the
    keyword arguments they pass are parameters of what they call"""
    import ast
    import importlib
    import inspect
    with open(DOC) as f:
        blocks = re.findall(r"```python\n(.*?)```", f.read(), re.S)
    assert len(blocks) >= 2
    checked = 0
    for b in blocks:
        tree = ast.parse(b)
        mods = {}
        for node in ast.walk(tree):
            if isinstance(node, ast.ImportFrom) and node.module == "ccfd_amd":
                for a in node.names:
                    mods[a.asname or a.name] = importlib.import_module(
                        "ccfd_amd." + a.name)
        for node in ast.walk(tree):
            if not (isinstance(node, ast.Attribute) and
                    isinstance(node.value, ast.Name) and node.value.id in mods):
                continue
            obj = getattr(mods[node.value.id], node.attr, None)
            assert obj is not None, (node.value.id, node.attr)
            checked += 1
        for node in ast.walk(tree):
            if not (isinstance(node, ast.Call) and
                    isinstance(node.func, ast.Attribute) and
                    isinstance(node.func.value, ast.Name) and
                    node.func.value.id in mods):
                continue
            fn = getattr(mods[node.func.value.id], node.func.attr)
            params = inspect.signature(fn).parameters
            if any(p.kind == p.VAR_KEYWORD for p in params.values()):
                continue
            for kw in node.keywords:
                assert kw.arg in params, (node.func.attr, kw.arg)
    assert checked >= 8
