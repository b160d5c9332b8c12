"""GPU parity of the image writers (cgs_image_warp = transformImage,
cgs_image_fuse = fuseAlignedImages / _local, cgs_write_png) against the
oracle's literal single-thread restatements (pinned in
tests/test_oracle_writers.py).  Bar: every element equal, including the
elements a writer leaves untouched and the local walk's overwrites."""
import io
import os
import subprocess

import numpy as np
import pytest

from cugridsearch_amd import api as A
from cugridsearch_amd import matcher as M
from cugridsearch_amd import synth
from cugridsearch_amd.build import DRIVER

pytestmark = pytest.mark.gpu


def _h(rng, scale=1.0):
    """A homography near a shift/rotation/zoom with a little perspective."""
    t = rng.uniform(-0.2, 0.2) * scale
    s = np.float32(1 + rng.uniform(-0.1, 0.1) * scale)
    h = [s * np.cos(t), -np.sin(t), rng.uniform(-6, 6), np.sin(t), s * np.cos(t), rng.uniform(-6, 6),
         rng.uniform(-2e-3, 2e-3) * scale, rng.uniform(-2e-3, 2e-3) * scale]
    return np.array(h, np.float32)


@pytest.mark.parametrize("in_t,out_t,c", [(np.uint8, np.uint8, 1), (np.uint8, np.uint8, 3), (np.float32, np.float32, 1),
                                          (np.float32, np.uint8, 1), (np.uint8, np.float32, 2)])
def test_warp_matches_oracle(gpu, oracle, in_t, out_t, c):
    rng = np.random.default_rng(11 + c)
    for k in range(12):
        ih, iw, oh, ow = (int(v) for v in rng.integers(5, 70, 4))
        src = rng.integers(0, 256, (ih, iw) + ((c,) if c > 1 else ())).astype(in_t)
        if in_t == np.float32:
            src = src + rng.random(src.shape).astype(np.float32)
        H = _h(rng)
        init = np.full((oh, ow) + ((c,) if c > 1 else ()), 7, out_t)   # untouched elements keep it
        want = oracle.image_warp(src, H, init.copy())
        out = A.CudaImage.from_numpy(init.copy())
        A.warp_image(A.CudaImage.from_numpy(src), H, ow, oh, out=out)
        got = out.tensor.cpu().numpy()
        assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), (k, ih, iw, oh, ow)


@pytest.mark.parametrize("local", [0, 1])
@pytest.mark.parametrize("dtype", [np.uint8, np.float32])
def test_fuse_matches_oracle(gpu, oracle, local, dtype):
    rng = np.random.default_rng(31 + local)
    # 23..29 wide: the local walk takes an extra step that wraps onto the next row
    for (h, w), (mh, mw) in [((20, 23), (20, 23)), ((24, 24), (30, 26)), ((17, 29), (17, 29)), ((9, 9), (12, 7)),
                             ((64, 48), (60, 50)), ((33, 25), (33, 25))]:
        ref = rng.integers(0, 256, (h, w)).astype(dtype)
        mov = rng.integers(0, 256, (mh, mw)).astype(dtype)
        est, truth = _h(rng, 0.5), _h(rng, 0.5)
        want = oracle.image_fuse(ref, mov, est, truth, 0, local, np.zeros((h, w, 3), dtype))
        got = A.fuse_images(A.CudaImage.from_numpy(ref), A.CudaImage.from_numpy(mov), est, truth, local=bool(local))
        assert np.array_equal(got.tensor.cpu().numpy().view(np.uint8), want.view(np.uint8)), (h, w, mh, mw)


def test_facade_writes_the_drivers_images(gpu, oracle, tmp_path):
    from PIL import Image
    if not os.path.exists(DRIVER):
        pytest.fail(f"{DRIVER} missing: run __graft_entry__.build()")
    m, f, _ = synth.make_pair(5, 40, 30, sigma=2.0, perturb=2.0, dtype=np.uint8)
    s, e, c = synth.grid_ranges(2, 2.0)
    spec = tmp_path / "spec.txt"
    spec.write_text("sad bilinear u8 40 30 1 0 8 " + " ".join(str(float(v)) for v in list(s) + list(e) + list(c)) + "\n")
    m.tofile(tmp_path / "m.raw")
    f.tofile(tmp_path / "f.raw")
    pre = str(tmp_path / "img")
    r = subprocess.run([DRIVER, str(spec), str(tmp_path / "m.raw"), str(tmp_path / "f.raw"), str(tmp_path / "c.bin"), pre],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    shift = np.array([1, 0, 2, 0, 1, 1, 0, 0], np.float32)
    ident = np.array([1, 0, 0, 0, 1, 0, 0, 0], np.float32)
    with Image.open(pre + "_out.png") as im:
        assert np.array_equal(np.asarray(im), oracle.image_warp(m, shift, np.zeros_like(m)))
    with Image.open(pre + "_fused.png") as im:
        assert np.array_equal(np.asarray(im), oracle.image_fuse(f, m, shift, ident, 0, 0, np.zeros((30, 40, 3), np.uint8)))
    with Image.open(pre + "_fused_local.png") as im:
        assert np.array_equal(np.asarray(im), oracle.image_fuse(f, m, ident, ident, 0, 1, np.zeros((30, 40, 3), np.uint8)))


def test_matcher_writes_output_images(gpu, oracle, tmp_path):
    from PIL import Image
    m, f, _ = synth.make_pair(21, 48, 40, sigma=2.0, perturb=2.0, dtype=np.uint8, gamma=True, noise=3.0)
    out_png, fused_png = str(tmp_path / "out.png"), str(tmp_path / "fused.png")
    res = M.run(f, m, out=io.StringIO(), log=io.StringIO(), depth=1, results_file=str(tmp_path / "r.txt"),
                out_name=out_png, fused_name=fused_png)
    H = np.array(res["homography"], np.float32)
    with Image.open(out_png) as im:
        got = np.asarray(im)
    assert np.array_equal(got, oracle.image_warp(m, H, np.zeros_like(f)))
    fix_filtered = res["fixed_filtered"].tensor.cpu().numpy()
    with Image.open(fused_png) as im:
        got = np.asarray(im)
    # no ground truth file: the truth channel uses the answer itself
    assert np.array_equal(got, oracle.image_fuse(fix_filtered, m, H, H, 0, 1, np.zeros((40, 48, 3), np.uint8)))
    M.write_image(str(tmp_path / "x.bmp"), res["warped"], out=(buf := io.StringIO()))
    assert "not recognized" in buf.getvalue() and not os.path.exists(tmp_path / "x.bmp")
