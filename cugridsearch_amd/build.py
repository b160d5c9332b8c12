"""Builds libcgs.so (every HIP source under csrc/, for gfx950) in-tree.

Flags that carry semantics:
  -ffp-contract=off  every + - * / is rounded on its own; FMA only where a
                     kernel writes fmaf()/fma() (the declared arithmetic the
                     exact search modes rely on);
  -target-feature -packed-fp32-ops   no packed FP32 instructions are emitted.
"""
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libcgs.so")
ARCH = os.environ.get("CGS_OFFLOAD_ARCH", "gfx950")


def hipcc_command(out=LIB, extra=()):
    srcs = sorted(glob.glob(os.path.join(HERE, "csrc", "*.hip")))
    return ["/opt/rocm/bin/hipcc", "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-shared",
            "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-command-line-argument", "-I", os.path.join(ROOT, "include"), "-I",
            os.path.join(HERE, "csrc"), *extra, "-o", out, *srcs]


def driver_command(out):
    """The C++ facade driver used by the GPU tests (tests/cpp/facade_driver.cpp)."""
    return ["/opt/rocm/bin/hipcc", "--offload-arch=" + ARCH, "-O2", "-std=c++17", "-Wno-unused-result",
            "-I", os.path.join(ROOT, "include"), "-o", out, os.path.join(ROOT, "tests", "cpp", "facade_driver.cpp"),
            "-L", HERE, "-lcgs", "-Wl,-rpath," + HERE, "-Wl,-rpath,$ORIGIN/../../../cugridsearch_amd"]


DRIVER = os.path.join(ROOT, "tests", "cpp", "_build", "facade_driver")


def build_driver(verbose=False):
    os.makedirs(os.path.dirname(DRIVER), exist_ok=True)
    cmd = driver_command(DRIVER)
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building the facade driver")
    return DRIVER


def build(verbose=False):
    cmd = hipcc_command(out=LIB + ".tmp")
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libcgs.so")
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
