"""A reference user's own __device__ functors through the C++ facade's generic
path (tests/cpp/functor_driver.cpp): bound through hipMemcpyFromSymbol as
the reference's drivers bind theirs, searched by value, by reference, by
workgroup (block / stream) and through search(); their costs and argmins
against the oracle and the built-in functor tag, the facade's device image
accessors against the oracle's samplers bit for bit, and the grid points the
functors receive against the oracle's decode."""
import json
import os
import subprocess

import numpy as np
import pytest

from cugridsearch_amd import synth
from cugridsearch_amd.build import FUNCTOR_DRIVER

pytestmark = pytest.mark.gpu

W, H = 40, 32
T_GRID = [(-3.0, 3.0, 7), (-2.0, 2.0, 5)]              # integer translations (x, y)
P_GRID = [(-1.5, W + 0.25, 87), (-1.25, H + 0.5, 69)]  # probe points (x, y), inside and out
S8 = [-0.1, 0.9, 0.95, -0.05, -4.0, -3.0, -1e-3, -2e-3]
E8 = [0.1, 1.1, 1.05, 0.05, 4.0, 3.0, 1e-3, 2e-3]
C8 = [3, 2, 3, 2, 5, 3, 2, 2]


@pytest.fixture(scope="module")
def driver_run(tmp_path_factory):
    if not os.path.exists(FUNCTOR_DRIVER):
        pytest.fail(f"{FUNCTOR_DRIVER} missing: run __graft_entry__.build()")
    d = tmp_path_factory.mktemp("functors")
    m, f, _ = synth.make_pair(91, W, H, sigma=2.0, perturb=2.0, dtype=np.uint8)
    spec = d / "spec.txt"
    rows = [f"{W} {H}"] + [" ".join(str(v) for v in g) for g in T_GRID] + \
        [" ".join(str(v) for v in g) for g in P_GRID] + [" ".join(str(v) for v in x) for x in (S8, E8, C8)]
    spec.write_text("\n".join(rows) + "\n")
    m.tofile(d / "m.raw")
    f.tofile(d / "f.raw")
    r = subprocess.run([FUNCTOR_DRIVER, str(spec), str(d / "m.raw"), str(d / "f.raw"), str(d)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    res = json.loads([l for l in r.stdout.splitlines() if l.startswith("{")][-1])
    out = {p[:-4]: np.fromfile(d / p, np.float32) for p in os.listdir(d) if p.endswith(".bin")}
    return m, f, res, out, r.stdout


def _grid2(oracle, g):
    s = np.float32([g[0][0], g[1][0]])
    e = np.float32([g[0][1], g[1][1]])
    c = np.float32([g[0][2], g[1][2]])
    c2, r = oracle.grid_update(s, e, c)
    return s, r, c2


def test_user_sad_every_launch_form(gpu, oracle, driver_run):
    """The user's SAD (integer translations, so every sum is exact) equals the
    oracle's sumOfAbsoluteDifferences bit for bit in every launch form, and so
    does the built-in tag; each argmin is the oracle's."""
    m, f, res, out, stdout = driver_run
    s, r, c2 = _grid2(oracle, T_GRID)
    want = oracle.search(m, f, oracle.COST_SAD_SUM, oracle.BILINEAR, s, r, c2, mode=oracle.PARAM_TRANSLATION)
    assert (want > 0).all()
    vo, io = oracle.argmin(want)
    for name in ("user_sad", "user_sad_search", "user_sad_ref", "user_sad_ref_stream", "user_sad_rows_stream",
                 "user_sad_rows_block", "tag_sad_sum"):
        assert np.array_equal(out[name].view(np.uint32), want.view(np.uint32)), name
        assert res[name][0] == io and np.float32(res[name][1]) == vo, (name, res[name], (vo, io))
    assert stdout.count("CudaGrid has") >= 7 and "Matrix translation_grid" in stdout


@pytest.mark.parametrize("sampler", ["bilinear", "nearest"])
def test_device_accessors_match_oracle_samplers(gpu, oracle, driver_run, sampler):
    """CudaImage::valueAt_bilinear / valueAt_nearest_neighbor in device code,
    at 6,003 points inside, on and outside the image (the last column's ceil
    tap wraps to the next row, the last row's reads 0), equal the oracle's
    samplers bit for bit."""
    m, f, res, out, _ = driver_run
    s, r, c2 = _grid2(oracle, P_GRID)
    got = out["probe_" + sampler]
    n = oracle.grid_num_elements(c2)
    assert got.shape == (n,)
    samp = oracle.BILINEAR if sampler == "bilinear" else oracle.NEAREST
    want = np.empty(n, np.float32)
    for k in range(n):
        x, y = oracle.grid_point(s, r, c2, k)
        want[k] = oracle.sample(m, y, x, samp)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), np.where(got != want)[0][:8]
    assert (want == 0).sum() > 100 and (want != 0).sum() > 3000


def test_functor_grid_points_match_oracle_decode(gpu, oracle, driver_run):
    """An 8-D grid: coordinate a of every point the functors receive equals
    the oracle's indexToGridPoint bit for bit."""
    _, _, res, out, _ = driver_run
    s, e, c = np.float32(S8), np.float32(E8), np.float32(C8)
    c2, r = oracle.grid_update(s, e, c)
    n = oracle.grid_num_elements(c2)
    pts = np.array([oracle.grid_point(s, r, c2, k) for k in range(n)], np.float32)
    for a in range(8):
        assert np.array_equal(out[f"axis{a}"].view(np.uint32), pts[:, a].view(np.uint32)), a
        assert res[f"axis{a}"][1] == pts[:, a].min()


def test_tensor_transform_and_indexing(gpu, driver_run):
    """CudaTensor::transform with a device lambda (2x + y on {3, 2} tensors),
    toIndexNd and toIndex1d with axis 0 fastest."""
    _, _, res, _, stdout = driver_run
    assert res["transform"] == [12.0, 72.0]
    assert res["toIndexNd5"] == [2, 1] and res["toIndex1d21"] == 5
    assert "Tensor A 3 x 2 elements of f" in stdout
