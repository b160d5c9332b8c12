"""GPU parity of the multispectral matcher: the image filter and quantisation
kernels bit-exact against the oracle, and the CLI run end to end on a small
PNG pair, its stdout parsed with the MATLAB harness's own patterns
(matlab/experiment_00.m:87-97) and its minimum checked against the oracle's
calcMIAlt search of the same grid."""
import io
import re

import numpy as np
import pytest

from cugridsearch_amd import api as A
from cugridsearch_amd import _lib as L
from cugridsearch_amd import matcher as M
from cugridsearch_amd import synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype,channels", [(np.uint8, 1), (np.uint8, 3), (np.float32, 1)])
def test_filter_matches_oracle(gpu, oracle, dtype, channels):
    rng = np.random.default_rng(3 + channels)
    shape = (37, 52) if channels == 1 else (37, 52, channels)
    img = (rng.random(shape) * 255).astype(dtype)
    filters = [M.UNSHARP_3X3, np.ones((1, 1), np.float32), np.full((5, 5), np.float32(1) / np.float32(25)),
               np.full((10, 10), np.float32(1) / np.float32(100)), rng.random((3, 4)).astype(np.float32) - 0.5]
    for filt in filters:
        got = A.filter_image(A.CudaImage.from_numpy(img), filt).tensor.cpu().numpy()
        want = oracle.image_filter(img, filt)
        assert got.dtype == want.dtype and np.array_equal(got.view(np.uint8), want.view(np.uint8)), filt.shape


def test_quantize_matches_oracle(gpu, oracle):
    v = np.tile(np.arange(256, dtype=np.uint8), 40).reshape(40, 256)
    for mod in (1.0, 2.5, 10.0, 0.3, 7.25, 100.0):
        img = A.CudaImage.from_numpy(v)
        A.quantize_image(img, mod)
        assert np.array_equal(img.tensor.cpu().numpy(), oracle.image_quantize(v, mod)), mod


def test_matcher_cli_protocol_and_minimum(gpu, oracle, tmp_path):
    from PIL import Image
    m, f, _ = synth.make_pair(21, 48, 40, sigma=2.0, perturb=2.0, dtype=np.uint8, gamma=True, noise=3.0)
    Image.fromarray(f).save(tmp_path / "fixed_0007.png")
    Image.fromarray(np.stack([m, m // 2, 255 - m], -1)).save(tmp_path / "moving.png")   # colour: stb's luma
    out = io.StringIO()
    res = M.run(M.load_gray(str(tmp_path / "fixed_0007.png")), M.load_gray(str(tmp_path / "moving.png")), out=out,
                results_file=str(tmp_path / "imageMatcherResults.txt"),
                names=(str(tmp_path / "fixed_0007.png"), str(tmp_path / "moving.png")))
    text = out.getvalue()
    # the harness: the first four {...} are the grid's start, end, step and count rows
    groups = re.findall(r"\{(.*?)\}", text, re.S)
    rows = [np.array([float(t) for t in gr.split(",")], np.float32) for gr in groups[:4]]
    num, start, end = M.coarse_grid(48, 40, 48, 40)
    assert np.allclose(rows[0], start, rtol=1e-5) and np.allclose(rows[1], end, rtol=1e-5)
    assert np.allclose(rows[3], num, rtol=1e-5)
    pat = r"Answer\[([^\]]*)]\s+Scale_factor\[([^\]]*)]\s+Homography_to_parameters\[([^\]]*)]\s+covar_matrix\[([^\]]*)]"
    mt = re.search(pat, text)
    assert mt, text[-800:]
    answer = np.array([float(t) for t in mt.group(1).split(",")])
    assert answer.shape == (8,) and np.allclose(answer, res["homography"], rtol=1e-5, atol=1e-6)
    assert [float(t) for t in mt.group(2).split(",")] == [1.0, 1.0]
    assert len(mt.group(3).split(",")) == 8 and len(mt.group(4).split(",")) == 64
    # the minimum: the oracle's calcMIAlt over the same grid (moving image filtered, fixed image not)
    mov = res["moving_filtered"].tensor.cpu().numpy()
    fix = M.load_gray(str(tmp_path / "fixed_0007.png"))
    assert np.array_equal(mov, oracle.image_filter(M.load_gray(str(tmp_path / "moving.png")), M.UNSHARP_3X3))
    c2, r = oracle.grid_update(start, end, num)
    want = oracle.search(mov, fix, oracle.COST_MI_ALT, oracle.BILINEAR, np.float32(start), r, c2,
                         mode=oracle.PARAM_NORM, bins=64)
    vo, io_ = oracle.argmin(want, oracle.TIE_REFERENCE)
    v, i = res["argmin"]
    assert abs(want[i] - vo) <= 2e-6 * abs(vo), (v, i, vo, io_)
    assert "Minimum found at point p = { " in text and "Covariance matrix:" in text
    line = (tmp_path / "imageMatcherResults.txt").read_text()
    assert line.startswith("input files,") and "Homography," in line and line.endswith("\n")
