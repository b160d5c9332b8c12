"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only).

* The oracle: its own CPU tests run again against an ASan/UBSan build of
  oracle/cgs_oracle.c (oracle/Makefile `sanitize`, loaded through
  CGS_ORACLE_SO with the sanitizer runtimes preloaded).
* libcgs: tests/cpp/host_sanitize.cpp drives the library's host paths (grid
  bookkeeping, sub-boxes, validation, the resize and local-fuse filter tables,
  the PNG reader on good, truncated, corrupted and oversized files) in a build whose host code is instrumented (-Xarch_host -fsanitize=...,
  libstdc++'s vector annotations on); without a GPU each device path stops
  at its first allocation.  The build takes over a minute, so it is kept
  under tests/cpp/_build and redone only when a source changes.

A sanitizer report, or a driver expectation that fails, fails the test."""
import glob
import hashlib
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "tests", "cpp", "_build")
GCC_RT = "/usr/lib/gcc/x86_64-linux-gnu/11"
REPORTS = ("ERROR: AddressSanitizer", "runtime error:", "ERROR: LeakSanitizer")


def _runtime(name):
    out = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
    return out if os.path.isabs(out) and os.path.exists(out) else None


def _san_env():
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=0"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1"
    return env


def test_oracle_under_sanitizers():
    asan, ubsan = _runtime("libasan.so"), _runtime("libubsan.so")
    if not asan or not ubsan:
        pytest.skip("gcc sanitizer runtimes not found")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True, capture_output=True)
    env = _san_env()
    env["CGS_ORACLE_SO"] = os.path.join(ROOT, "oracle", "_build", "san", "libcgs_oracle.so")
    env["LD_PRELOAD"] = asan + ":" + ubsan
    env["OMP_NUM_THREADS"] = "4"
    files = ["test_oracle.py", "test_oracle_alt.py", "test_oracle_writers.py", "test_oracle_resize.py",
             "test_matcher.py", "test_oracle_tensor.py"]
    code = ("import sys, pytest; from oracle import oracle as O; O.lib(); "
            "assert O._SO == sys.argv[1], O._SO; sys.exit(pytest.main(sys.argv[2:]))")
    r = subprocess.run([sys.executable, "-c", code, env["CGS_ORACLE_SO"], "-x", "-q", "-p", "no:cacheprovider",
                        *[os.path.join(ROOT, "tests", f) for f in files]],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    log = r.stdout + r.stderr
    assert r.returncode == 0, log[-4000:]
    for s in REPORTS:
        assert s not in log, log[-4000:]


def _sources():
    return sorted(glob.glob(os.path.join(ROOT, "cugridsearch_amd", "csrc", "*"))) + \
        sorted(glob.glob(os.path.join(ROOT, "include", "*.h"))) + [os.path.join(ROOT, "tests", "cpp", "host_sanitize.cpp")]


def _build_driver():
    exe = os.path.join(BUILD, "host_sanitize")
    h = hashlib.sha256()
    for p in _sources():
        h.update(p.encode())
        h.update(open(p, "rb").read())
    stamp = exe + ".sha256"
    if os.path.exists(exe) and os.path.exists(stamp) and open(stamp).read() == h.hexdigest():
        return exe
    os.makedirs(BUILD, exist_ok=True)
    cmd = ["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O1", "-g", "-std=c++17", "-ffp-contract=off",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
           "-D_GLIBCXX_SANITIZE_VECTOR=1", "-fno-omit-frame-pointer", "-Wno-unused-result",
           "-Wno-unused-command-line-argument", "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(ROOT, "cugridsearch_amd", "csrc"), "-o", exe + ".tmp",
           os.path.join(ROOT, "tests", "cpp", "host_sanitize.cpp"),
           *sorted(glob.glob(os.path.join(ROOT, "cugridsearch_amd", "csrc", "*.hip"))), "-lz"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, (r.stdout + r.stderr)[-4000:]
    os.replace(exe + ".tmp", exe)
    open(stamp, "w").write(h.hexdigest())
    return exe


def test_libcgs_host_code_under_sanitizers():
    if not os.path.exists("/opt/rocm/bin/hipcc"):
        pytest.skip("hipcc not found")
    exe = _build_driver()
    env = _san_env()
    env["HIP_VISIBLE_DEVICES"] = ""   # host paths only, also where a GPU is present
    r = subprocess.run([exe], env=env, capture_output=True, text=True, timeout=600)
    log = r.stdout + r.stderr
    assert r.returncode == 0 and "host_sanitize ok" in log, log[-4000:]
    for s in REPORTS:
        assert s not in log, log[-4000:]
