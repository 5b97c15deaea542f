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
GPU_SRC = ["ccfd_gpu.hip"]

# FP64 parity needs IEEE division/sqrt and no contraction (the reference is
# built with -std=c99, i.e. -ffp-contract=off). Packed-FP32 generation is off.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math", "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
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
    if os.path.exists(gpu):
        cmd += ["-L" + LIB, "-lccfd_gpu", "-Wl,-rpath,$ORIGIN"]
    _run(cmd)
    return out


def build_gpu():
    os.makedirs(LIB, exist_ok=True)
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    srcs = [os.path.join(CSRC, s) for s in GPU_SRC
            if os.path.exists(os.path.join(CSRC, s))]
    if not srcs:
        return None
    out = os.path.join(LIB, "libccfd_gpu.so")
    tmp = out + ".tmp"
    _run([hipcc] + HIP_FLAGS + ["-I" + INC, "-I" + CSRC, "-o", tmp] + srcs)
    os.replace(tmp, out)
    return out


def build_oracle():
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
