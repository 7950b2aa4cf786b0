"""Builds libcfs.so (HIP kernels + C-ABI) in-tree for gfx950 with hipcc.

Usage: python 2d_compressible_flow_solver_amd/build.py [-v]

Flags that matter for results:
  -ffp-contract=off   no FMA contraction: FP64 kernels must reproduce the
                      reference's separate multiply/add roundings;
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed FP32 instructions (they have given wrong
                      results on these machines).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcfs.so")
RUN = os.path.join(HERE, "cfs_run")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function"]
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), r.stdout))
    if verbose and out:
        print(out)


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    jobs = []
    for src in ["cfs_api.hip", "batch.hip", "ib_setup.hip"]:   # compiled in parallel
        obj = os.path.join(BUILD, src + ".o")
        cmd = [HIPCC, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj] + COMMON + DEVICE
        if verbose:
            print(" ".join(cmd), flush=True)
        jobs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        objs.append(obj)
    for cmd, pr in jobs:
        out, _ = pr.communicate()
        if pr.returncode != 0:
            raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), out))
    for src in ["serve.cpp"]:
        obj = os.path.join(BUILD, src + ".o")
        _run(["g++", "-c", os.path.join(CSRC, src), "-o", obj] + COMMON, verbose)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-o", tmp] + objs + ["--offload-arch=" + ARCH, "-fPIC"], verbose)
    os.replace(tmp, LIB)
    # the driver: the reference's Run / param.h / plugin roles on the C-ABI
    drv = os.path.join(HERE, "driver")
    tmp = RUN + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-Wall", "-o", tmp,
          os.path.join(drv, "run.cpp"), os.path.join(drv, "main.cpp"),
          "-L" + HERE, "-lcfs", "-ldl", "-Wl,-rpath,$ORIGIN"], verbose)
    os.replace(tmp, RUN)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
