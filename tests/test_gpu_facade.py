"""The C++ facade (include/cuGridSearch.hpp) driven the way the reference's
drivers drive its headers: a compiled driver (tests/cpp/facade_driver.cpp)
builds CudaImage / CudaGrid / CudaTensor / CudaGridSearcher, searches, calls
find_extrema and getGridPoint; its costs and argmins are compared with the
oracle."""
import json
import os
import subprocess

import numpy as np
import pytest

from cugridsearch_amd import synth
from cugridsearch_amd.build import DRIVER

pytestmark = pytest.mark.gpu


def run_driver(tmp_path, fn, sampler, moved, fixed, start, end, count, tie=0, bins=32):
    if not os.path.exists(DRIVER):
        pytest.fail(f"{DRIVER} missing: run __graft_entry__.build()")
    pix = "u8" if moved.dtype == np.uint8 else "f32"
    h, w = fixed.shape
    spec = tmp_path / "spec.txt"
    vals = " ".join(str(float(v)) for v in list(start) + list(end) + list(count))
    spec.write_text(f"{fn} {sampler} {pix} {w} {h} 1 {tie} {bins} {vals}\n")
    mp, fp, out = tmp_path / "m.raw", tmp_path / "f.raw", tmp_path / "costs.bin"
    moved.tofile(mp)
    fixed.tofile(fp)
    r = subprocess.run([DRIVER, str(spec), str(mp), str(fp), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    line = [l for l in r.stdout.splitlines() if l.startswith("{")][-1]
    assert "CudaGrid has" in r.stdout and "Grid MIN value has idx" in r.stdout
    return json.loads(line), np.fromfile(out, np.float32), r.stdout


def test_facade_nn_sad_bit_exact(gpu, oracle, tmp_path):
    m, f, s, e, c, _ = synth.workload("nn_sad_64", levels=3)
    res, costs, out = run_driver(tmp_path, "sad", "nearest", m, f, s, e, c)
    c2, r = oracle.grid_update(s, e, c)
    want = oracle.search(m, f, oracle.COST_SAD, oracle.NEAREST, s, r, c2)
    assert np.array_equal(costs.view(np.uint32), want.view(np.uint32))
    v, i = oracle.argmin(want)
    assert res["fused_index"] == res["extrema_index"] == i
    assert res["point0"] == oracle.grid_point(s, r, c2, i)[0]
    assert out.count("{") >= 4      # grid.display(): the four rows the MATLAB harness parses


def test_facade_ncc_and_reference_ties(gpu, oracle, tmp_path):
    m, f, _ = synth.make_pair(77, 48, sigma=3.0, perturb=3.0, dtype=np.float32)
    s, e, c = synth.grid_ranges(3, 3.0)
    res, costs, _ = run_driver(tmp_path, "ncc", "bilinear", m, f, s, e, c, tie=1)
    c2, r = oracle.grid_update(s, e, c)
    want = oracle.search(m, f, oracle.COST_NCC, oracle.BILINEAR, s, r, c2)
    assert np.allclose(costs, want, rtol=1e-5, atol=0)
    vo, io = oracle.argmin(want, oracle.TIE_REFERENCE)
    assert abs(want[res["fused_index"]] - vo) <= 2e-5 * abs(vo)
    assert res["fused_index"] == res["extrema_index"]
    cov, _, _ = oracle.covariance(s, r, c2, costs)      # covariance_by_value over the same tensor
    assert abs(res["covar_trace"] - np.trace(cov)) <= 1e-5 * abs(np.trace(cov))


def test_facade_mi(gpu, oracle, tmp_path):
    m, f, _ = synth.make_pair(78, 56, sigma=3.0, perturb=3.0, dtype=np.uint8, gamma=True, noise=4.0)
    s, e, c = synth.grid_ranges(3, 3.0)
    res, costs, _ = run_driver(tmp_path, "mi", "bilinear", m, f, s, e, c, bins=32)
    c2, r = oracle.grid_update(s, e, c)
    want = oracle.search(m, f, oracle.COST_MI, oracle.BILINEAR, s, r, c2, bins=32)
    assert np.allclose(costs, want, rtol=1e-6, atol=0)
    assert res["fused_index"] == res["extrema_index"]
