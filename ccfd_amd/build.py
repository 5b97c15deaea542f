"""Build the native libraries in-tree (no JIT caches, so the .so files travel
with the source tree).

  ccfd_amd/lib/libccfd_host.so  - C host library (mesh preprocessing, ini
                                  parser, drop-in driver), gcc -std=c99
  ccfd_amd/lib/libccfd_gpu.so   - HIP kernels + C ABI, hipcc for gfx950
  oracle/libccfd_oracle.so      - CPU restatement (test infrastructure)
  oracle/_ref/*.so              - the reference's operator sources + harness,
                                  only where the reference tree is present
"""
import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "ccfd_amd")
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "lib")
INC = os.path.join(ROOT, "include")
ORACLE = os.path.join(ROOT, "oracle")

HOST_SRC = ["mesh.c", "ini.c", "driver.c"]
GPU_SRC = ["ccfd_gpu.hip", "implicit.hip", "batch.hip", "analyze.hip",
           "devmesh.hip"]
HOST_CXX_SRC = ["soa.cpp", "agglomerate.cpp", "exact_host.cpp", "serve.cpp"]   # host C++ in libccfd_gpu.so
FLUX_CODES = list(range(1, 12))
OBJ = os.path.join(ROOT, "_build")

# FP64 parity needs IEEE division/sqrt and no contraction (the reference is
# built with -std=c99, i.e. -ffp-contract=off). Packed-FP32 generation is off.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", "-Wno-unused-value",
    "-Wno-unused-result",
]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build failed: " + " ".join(cmd))
    return r.stdout


def build_host():
    os.makedirs(LIB, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in HOST_SRC
            if os.path.exists(os.path.join(CSRC, s))]
    out = os.path.join(LIB, "libccfd_host.so")
    gpu = os.path.join(LIB, "libccfd_gpu.so")
    cmd = ["gcc", "-std=c99", "-D_POSIX_C_SOURCE=200809L", "-O2",
           "-ffp-contract=off", "-fPIC", "-shared", "-Wall", "-I" + INC,
           "-o", out] + srcs + ["-lm"]
    if not os.path.exists(gpu):
        raise RuntimeError("build the GPU library first (libccfd_gpu.so)")
    cmd += ["-L" + LIB, "-lccfd_gpu", "-Wl,-rpath,$ORIGIN", "-ldl"]
    _run(cmd)
    build_cgns()
    # `ccfd case.ini`: the reference's command line on the GPU path
    _run(["gcc", "-std=c99", "-O2", "-I" + INC, "-o", os.path.join(LIB, "ccfd"),
          os.path.join(CSRC, "ccfd_main.c"), "-L" + LIB, "-lccfd_host",
          "-lccfd_gpu", "-Wl,-rpath,$ORIGIN", "-lm"])
    return out


def hdf5_prefix():
    """an HDF5 installation with the C headers and library, or None"""
    import glob
    for pre in (os.environ.get("HDF5_DIR"), "/opt/conda", "/usr",
                "/usr/lib/x86_64-linux-gnu/hdf5/serial"):
        if not pre:
            continue
        inc = os.path.join(pre, "include", "hdf5.h")
        libs = glob.glob(os.path.join(pre, "lib", "libhdf5.so*"))
        if os.path.exists(inc) and libs:
            return pre
    return None


def build_cgns():
    """libccfd_cgns.so: the CGNS output and restart files of the drop-in,
    written with HDF5 (loaded by the drop-in only when a case uses CGNS).
    Without HDF5 it is not built, and such cases stop with a message."""
    out = os.path.join(LIB, "libccfd_cgns.so")
    pre = hdf5_prefix()
    if pre is None:
        if os.path.exists(out):
            os.remove(out)
        return None
    _run(["gcc", "-std=c99", "-D_POSIX_C_SOURCE=200809L", "-O2", "-fPIC",
          "-shared", "-Wall", "-I" + INC, "-I" + os.path.join(pre, "include"),
          "-o", out, os.path.join(CSRC, "cgns_hdf5.c"),
          "-L" + os.path.join(pre, "lib"), "-lhdf5",
          "-Wl,-rpath," + os.path.join(pre, "lib")])
    return out


def build_gpu(jobs=None):
    """Compile every HIP object for gfx950 (in parallel) and link the C ABI
    library. One object per Riemann flux keeps compile times short."""
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(OBJ, exist_ok=True)
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    inc = ["-I" + INC, "-I" + CSRC]
    units = []
    for s in GPU_SRC:
        units.append((os.path.join(CSRC, s),
                      os.path.join(OBJ, s.replace(".hip", ".o")), []))
    for k in FLUX_CODES:
        units.append((os.path.join(CSRC, "op_flux.hip"),
                      os.path.join(OBJ, "op_flux_%d.o" % k),
                      ["-DFLUX_CODE=%d" % k]))

    for s in HOST_CXX_SRC:
        units.append((os.path.join(CSRC, s),
                      os.path.join(OBJ, s.replace(".cpp", ".o")), None))

    def comp(u):
        src, obj, extra = u
        if extra is None:   # host-only C++
            _run(["g++", "-std=c++17", "-O2", "-fPIC", "-ffp-contract=off"] +
                 inc + ["-c", src, "-o", obj])
        else:
            _run([hipcc] + HIP_FLAGS + inc + extra + ["-c", src, "-o", obj])
        return obj

    jobs = jobs or min(8, os.cpu_count() or 4)
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(comp, units))
    out = os.path.join(LIB, "libccfd_gpu.so")
    tmp = out + ".tmp"
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, out)
    return out


def build_test_native():
    """host builds for the CPU tests (test-only). -Bsymbolic: each harness
    calls the copy of the code compiled into it, not the same-named symbol
    of libccfd_gpu.so when that is loaded in the same process."""
    d = os.path.join(ROOT, "tests", "native")
    src = os.path.join(d, "physics_host.cpp")
    if os.path.exists(src):
        _run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-fPIC",
              "-shared", "-I" + CSRC, "-o",
              os.path.join(d, "libphysics_host.so"), src])
    src = os.path.join(d, "agglomerate_host.cpp")
    if os.path.exists(src):
        _run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-I" + CSRC,
              "-Wl,-Bsymbolic", "-o",
              os.path.join(d, "libagglomerate_host.so"), src,
              os.path.join(CSRC, "agglomerate.cpp")])
    src = os.path.join(d, "soa_host.cpp")
    if os.path.exists(src):
        _run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-I" + CSRC,
              "-I" + INC, "-Wl,-Bsymbolic", "-o",
              os.path.join(d, "libsoa_host.so"), src,
              os.path.join(CSRC, "soa.cpp")])
    src = os.path.join(d, "range_alloc_host.cpp")
    if os.path.exists(src):
        _run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-I" + CSRC,
              "-o", os.path.join(d, "librange_alloc_host.so"), src])


def build_oracle():
    build_test_native()
    _run(["make", "-s", "-C", ORACLE, "oracle"])
    if os.path.exists("/root/reference/src/finiteVolume.c"):
        _run(["make", "-s", "-C", ORACLE, "ref"])


def build_all():
    build_gpu()
    build_host()
    build_oracle()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "host": build_host, "gpu": build_gpu,
         "oracle": build_oracle}[w]()
    print("built:", ", ".join(what))
