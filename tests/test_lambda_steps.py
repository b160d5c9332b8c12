"""The closed-form lambda sequences of the line-traversal kernels
(csrc/cgs_lambda.hpp) against the literal binary32 loops, and the binary32
form of calcMIAlt's power-of-two rescale against its binary64 expression,
compiled for the host with g++ -ffp-contract=off.  CPU only."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def test_closed_form_lambda_matches_sequential_loops(tmp_path):
    exe = str(tmp_path / "lambda_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
                    "-I", os.path.join(ROOT, "cugridsearch_amd", "csrc"), "-o", exe,
                    os.path.join(HERE, "cpp", "lambda_check.cpp")], check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
    bad, n, runs512, direct, chains, rescales, areas, fastpass = (int(v) for v in out[:8])
    assert n > 2_000_000 and bad == 0, (bad, n)
    # the direct form saw the integer starts and the starts from 16 up; the
    # chain every start below 2^23
    assert direct > 40000 and chains > 80000, (direct, chains)
    # calcMIAlt's binary32 rescale for power-of-two sizes equals the binary64
    # expression on every compared value
    assert rescales > 8_000_000, rescales
    # a 512-wide row (step 1/511, a tie in one binade) from lambda = 0 (twice:
    # +0 and -0) takes its 512 steps in 8 full passes of 64, its short
    # segments near 0 sharing the first
    assert runs512 == 2 * 8, runs512
    # most passes of the walks are whole-pass closed forms
    assert fastpass > 100_000, fastpass
    # the moved area's closed-form repeated add equals the literal loop
    assert areas > 100_000, areas
