"""Builds the native library libmhdg.so in-tree (mpcns_uns_amd/_lib/).

Host setup code is compiled with g++ (OpenMP), the HIP kernels and the device context with hipcc for
gfx950; both without floating-point contraction, which the bit-exact parity with the reference's CPU
arithmetic relies on.
"""
import glob
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_lib")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = os.environ.get("MHDG_ARCH", "gfx950")

HOST_SRC = sorted(os.path.relpath(p, CSRC) for p in glob.glob(os.path.join(CSRC, "host", "*.cpp")))
HIP_SRC = sorted(os.path.relpath(p, CSRC) for p in glob.glob(os.path.join(CSRC, "kernels", "*.hip")))
FP = ["-ffp-contract=off", "-fno-fast-math"]
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", *FP,
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-munsafe-fp-atomics"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + os.path.basename(cmd[-1]))
    return r


def build(verbose=False):
    os.makedirs(os.path.join(OUT, "obj"), exist_ok=True)
    jobs = []
    for s in HOST_SRC:
        o = os.path.join(OUT, "obj", os.path.basename(s) + ".o")
        jobs.append(["g++", "-std=c++17", "-O2", "-fPIC", "-fopenmp", *FP, "-c", os.path.join(CSRC, s), "-o", o])
    for s in HIP_SRC:
        o = os.path.join(OUT, "obj", os.path.basename(s) + ".o")
        jobs.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, s), "-o", o])
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        list(ex.map(_run, jobs))
    objs = [j[-1] for j in jobs]
    lib = os.path.join(OUT, "libmhdg.so")
    tmp = lib + ".tmp"
    _run(["g++", "-shared", "-fopenmp", "-o", tmp, *objs, "-L" + os.path.join(ROCM, "lib"),
          "-Wl,-rpath," + os.path.join(ROCM, "lib"), "-lamdhip64"])
    os.replace(tmp, lib)
    if verbose:
        print("built", lib)
    return lib


if __name__ == "__main__":
    build(verbose=True)
