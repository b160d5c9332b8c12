"""The MEX gateways (cugridsearch_amd/mex/) through the MEX API stand-in of
tests/mex/, on the GPU: cgs_mex (the reference's cudaAdd,
src/gpu/cudaGridSearch_mex.cu:48-114) element for element against the host
sum, and cgs_search_mex (column-major MATLAB images in, the argmin's point,
cost, index and every cost out) against the oracle."""
import json
import os
import subprocess

import numpy as np
import pytest

from cugridsearch_amd import synth
from cugridsearch_amd.build import MEX_CHECKS

pytestmark = pytest.mark.gpu


def _exe(gateway):
    exe = MEX_CHECKS[gateway]
    if not os.path.exists(exe):
        pytest.fail(f"{exe} missing: run __graft_entry__.build()")
    return exe


def test_mex_add(gpu):
    r = subprocess.run([_exe("cgs_mex.c")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "mex_add ok" in r.stdout, r.stdout + r.stderr
    assert "-1.250000" in r.stdout and "1.000000" in r.stdout   # the gateway's prints of A[0], B[1]


CASES = [
    # functor, sampler, dtype, side (W, H), grid (levels, translation), bar
    ("centredAbsoluteDifference", "nearest", np.uint8, (48, 40), (3, 3.0), "exact"),
    ("calcNCC", "bilinear", np.float32, (56, 44), (3, 3.0), 1e-5),
    ("calcMIstream", "bilinear", np.uint8, (48, 48), (3, 3.0), 1e-6),
    ("calcNCCAlt", "bilinear", np.float32, (40, 36), (3, 3.0), "exact"),
]


@pytest.mark.parametrize("functor,sampler,dtype,size,grid,bar", CASES, ids=[c[0] for c in CASES])
def test_mex_search_matches_oracle(gpu, oracle, tmp_path, functor, sampler, dtype, size, grid, bar):
    W, H = size
    m, f, _ = synth.make_pair(61, W, H, sigma=2.0, perturb=2.5, dtype=dtype)
    s, e, c = synth.grid_ranges(*grid)
    bins = 32
    spec = tmp_path / "spec.txt"
    vals = " ".join(str(float(v)) for v in list(s) + list(e) + list(c))
    spec.write_text(f"{'u8' if dtype == np.uint8 else 'f32'} {W} {H} 8 {vals} {functor} {bins} {sampler}\n")
    m.tofile(tmp_path / "m.raw")
    f.tofile(tmp_path / "f.raw")
    r = subprocess.run([_exe("cgs_search_mex.c"), str(spec), str(tmp_path / "m.raw"), str(tmp_path / "f.raw"),
                        str(tmp_path / "costs.bin")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    res = json.loads([l for l in r.stdout.splitlines() if l.startswith("{")][-1])
    got = np.fromfile(tmp_path / "costs.bin", np.float32)
    from cugridsearch_amd import api as A
    fn = getattr(A, functor)
    c2, rr = oracle.grid_update(s, e, c)
    want = oracle.search(m, f, fn.cost, oracle.NEAREST if sampler == "nearest" else oracle.BILINEAR, s, rr, c2,
                         mode=fn.param_mode, bins=bins)
    assert got.shape == want.shape == (res["n"],)
    vo, io = oracle.argmin(want)
    if bar == "exact":
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
        assert res["index"] == io and np.float32(res["cost"]) == vo
    else:
        assert np.all(np.abs(got - want) <= bar * np.abs(want))
        assert abs(want[int(res["index"])] - vo) <= 2 * bar * abs(vo)
    assert np.array_equal(np.float32(res["point"]), oracle.grid_point(s, rr, c2, int(res["index"])))
