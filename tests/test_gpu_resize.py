"""GPU parity of cgs_image_resize (stbir_resize_uint8 as the reference's
matcher calls it, src/cpu/multispectral_ImageMatcher.cu:531, 550) against the
oracle's restatement, which tests/test_oracle_resize.py pins bit for bit to
the library itself.  Bar: every byte equal."""
import io
import re

import numpy as np
import pytest

from cugridsearch_amd import api as A
from cugridsearch_amd import _lib as L
from cugridsearch_amd import matcher as M
from cugridsearch_amd import synth

pytestmark = pytest.mark.gpu


def _image(rng, h, w, c, smooth):
    a = rng.integers(0, 256, (h, w, c) if c > 1 else (h, w), dtype=np.uint8)
    if smooth:
        a = (np.cumsum(a.astype(np.int64), axis=1) // max(w, 1) % 256).astype(np.uint8)
    return a


def test_random_sizes_bit_exact(gpu, oracle):
    rng = np.random.default_rng(77)
    for k in range(80):
        iw, ih, ow, oh = (int(v) for v in rng.integers(1, 130, 4))
        c = int(rng.integers(1, 5))
        a = _image(rng, ih, iw, c, k % 2 == 0)
        got = A.resize_image(A.CudaImage.from_numpy(a), ow, oh).tensor.cpu().numpy()
        assert np.array_equal(got, oracle.image_resize(a, ow, oh)), (iw, ih, ow, oh, c)


@pytest.mark.parametrize("shape", [((700, 3), (2, 5)), ((1, 1), (64, 33)), ((64, 33), (1, 1)), ((97, 61), (97, 61)),
                                   ((512, 384), (256, 256)), ((256, 256), (512, 384)), ((300, 100), (100, 200))],
                         ids=lambda s: f"{s[0][0]}x{s[0][1]}->{s[1][0]}x{s[1][1]}")
def test_extreme_and_matcher_shapes(gpu, oracle, shape):
    (iw, ih), (ow, oh) = shape
    rng = np.random.default_rng(iw + 7 * ih)
    a = _image(rng, ih, iw, 1, True)
    got = A.resize_image(A.CudaImage.from_numpy(a), ow, oh).tensor.cpu().numpy()
    assert np.array_equal(got, oracle.image_resize(a, ow, oh))


def test_resize_rejects_bad_arguments(gpu):
    u8 = A.CudaImage.from_numpy(np.zeros((8, 8), np.uint8))
    f32 = A.CudaImage.from_numpy(np.zeros((8, 8), np.float32))
    with pytest.raises(L.CgsError):
        A.resize_image(f32, 4, 4)   # uint8 only, as stbir_resize_uint8
    rgb = A.CudaImage.from_numpy(np.zeros((4, 4, 3), np.uint8))
    import ctypes
    st = L.lib().cgs_image_resize(ctypes.byref(u8._img), ctypes.byref(rgb._img), None)
    assert st != 0 and "channel" in L.lib().cgs_last_error().decode()
    st = L.lib().cgs_image_resize(ctypes.byref(u8._img), ctypes.byref(u8._img), None)
    assert st != 0 and "alias" in L.lib().cgs_last_error().decode()


def test_matcher_resizes_the_larger_image(gpu, oracle, tmp_path):
    """Fixed 72x60 against moving 48x40 (2.25x the area): the fixed image is
    brought to 48x40 with the library's resampler, the scale factors are
    2/3, and the search runs on the resized pair."""
    m, f, _ = synth.make_pair(21, 48, 40, sigma=2.0, perturb=2.0, dtype=np.uint8, gamma=True, noise=3.0)
    big = oracle.image_resize(f, 72, 60)   # any 72x60 image will do
    out, log = io.StringIO(), io.StringIO()
    res = M.run(big, m, out=out, log=log, depth=1, results_file=str(tmp_path / "r.txt"),
                names=("fixed_0003.png", "moving.png"))
    assert "Rescaling fixed_0003.png from (72,60) to (48,40)" in log.getvalue()
    name, img = res["resized"]
    assert name == "image_fixed_resized.png"
    assert np.array_equal(img.tensor.cpu().numpy(), oracle.image_resize(big, 48, 40))
    sf = np.float32(48) / np.float32(72)
    assert res["scale_factor"] == (sf, np.float32(40) / np.float32(60))
    mt = re.search(r"Scale_factor\[([^\]]*)]", out.getvalue())
    assert mt and [float(t) for t in mt.group(1).split(",")] == pytest.approx([0.666667, 0.666667], rel=1e-6)
