"""Builds libcgs.so (every HIP source under csrc/, for gfx950) in-tree.

Flags that carry semantics:
  -ffp-contract=off  every + - * / is rounded on its own; FMA only where a
                     kernel writes fmaf()/fma() (the declared arithmetic the
                     exact search modes rely on);
  -target-feature -packed-fp32-ops   no packed FP32 instructions are emitted.
"""
import glob
import hashlib
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libcgs.so")
ARCH = os.environ.get("CGS_OFFLOAD_ARCH", "gfx950")


def source_files():
    """The files libcgs.so is built from: csrc/* and include/*."""
    return sorted(glob.glob(os.path.join(HERE, "csrc", "*"))) + sorted(glob.glob(os.path.join(ROOT, "include", "*")))


def source_hash():
    """First 16 hex digits of the SHA-256 of the library's source files (path
    relative to the repository, then content, in sorted order).  build()
    embeds it in libcgs.so, where cgs_version() reports it."""
    h = hashlib.sha256()
    for p in source_files():
        h.update(os.path.relpath(p, ROOT).encode())
        with open(p, "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def hipcc_command(out=LIB, extra=()):
    srcs = sorted(glob.glob(os.path.join(HERE, "csrc", "*.hip")))
    return ["/opt/rocm/bin/hipcc", "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-shared",
            "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-command-line-argument", f'-DCGS_SOURCE_HASH="{source_hash()}"',
            "-I", os.path.join(ROOT, "include"), "-I", os.path.join(HERE, "csrc"), *extra, "-o", out, *srcs, "-lz"]


def driver_command(out, source="facade_driver.cpp"):
    """A C++ driver of the facade used by the GPU tests (tests/cpp/<source>)."""
    return ["/opt/rocm/bin/hipcc", "--offload-arch=" + ARCH, "-O2", "-std=c++17", "-Wno-unused-result",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument",
            "-I", os.path.join(ROOT, "include"), "-o", out, os.path.join(ROOT, "tests", "cpp", source),
            "-L", HERE, "-lcgs", "-Wl,-rpath," + HERE, "-Wl,-rpath,$ORIGIN/../../../cugridsearch_amd"]


DRIVER = os.path.join(ROOT, "tests", "cpp", "_build", "facade_driver")
FUNCTOR_DRIVER = os.path.join(ROOT, "tests", "cpp", "_build", "functor_driver")


def build_driver(verbose=False):
    """Builds both facade drivers: the tag-functor driver and the one with
    user-written __device__ functors."""
    os.makedirs(os.path.dirname(DRIVER), exist_ok=True)
    for out, src in ((DRIVER, "facade_driver.cpp"), (FUNCTOR_DRIVER, "functor_driver.cpp")):
        cmd = driver_command(out, src)
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("hipcc failed building " + src)
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


MEX_DIR = os.path.join(HERE, "mex")
MEX_CHECKS = {"cgs_mex.c": os.path.join(ROOT, "tests", "mex", "_build", "mex_add_check"),
              "cgs_search_mex.c": os.path.join(ROOT, "tests", "mex", "_build", "mex_search_check")}


def mex_command(gateway, out):
    """A MEX gateway of cugridsearch_amd/mex/ linked with the MEX API stand-in
    of tests/mex/ and its check program, as a plain executable (no MATLAB)."""
    check = os.path.join(ROOT, "tests", "mex", os.path.basename(out) + ".c")
    return ["gcc", "-O2", "-std=c11", "-Wall", "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROOT, "tests", "mex"),
            "-I", os.path.join(ROOT, "include"), "-I", "/opt/rocm/include", "-o", out,
            os.path.join(MEX_DIR, gateway), os.path.join(ROOT, "tests", "mex", "fake_mex.c"), check,
            "-L", HERE, "-lcgs", "-L", "/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath," + HERE,
            "-Wl,-rpath,/opt/rocm/lib"]


def build_mex(verbose=False):
    for gateway, out in MEX_CHECKS.items():
        os.makedirs(os.path.dirname(out), exist_ok=True)
        cmd = mex_command(gateway, out)
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("gcc failed building the MEX check of " + gateway)


CLI = os.path.join(HERE, "_build", "cgs_matcher")


def cli_command(out=CLI):
    """The matcher CLI (cli/cgs_matcher.cpp): host C++ over the C-ABI, linked
    against libcgs.so (found next to it through the rpath) and the HIP
    runtime.  -ffp-contract=off keeps its binary32 expressions unfused."""
    return ["g++", "-O2", "-std=c++17", "-Wall", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__",
            "-I", os.path.join(ROOT, "include"), "-I", "/opt/rocm/include", "-o", out,
            os.path.join(HERE, "cli", "cgs_matcher.cpp"), "-L", HERE, "-lcgs", "-L", "/opt/rocm/lib", "-lamdhip64",
            "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath,/opt/rocm/lib"]


def build_cli(verbose=False):
    os.makedirs(os.path.dirname(CLI), exist_ok=True)
    cmd = cli_command(CLI + ".tmp")
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("g++ failed building cgs_matcher")
    os.replace(CLI + ".tmp", CLI)
    return CLI


if __name__ == "__main__":
    print(build(verbose=True))
