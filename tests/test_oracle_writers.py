"""The image-writer restatements of the oracle (transformImage,
fuseAlignedImages, fuseAlignedImages_local; oracle/cgs_oracle.c) pinned by
known answers that a wrong sign, a transposed operand or a swapped channel
fails: the identity copies, an integer shift moves the image and leaves the
rest untouched, a quarter turn is numpy's rot90, the overlay's channels hold
the right image under the right homography, and the normalised-coordinate
walk lands every power-of-two step exactly on its pixel.  Also the PNG writer
of the library (host code, no GPU call), read back with PIL."""
import numpy as np
import pytest

from oracle import oracle as O

I8 = np.array([1, 0, 0, 0, 1, 0, 0, 0], np.float32)


def _img(rng, h, w, c=1, dtype=np.uint8):
    shape = (h, w) if c == 1 else (h, w, c)
    a = rng.integers(0, 256, shape)
    return a.astype(dtype)


@pytest.mark.parametrize("dtype,c", [(np.uint8, 1), (np.uint8, 3), (np.float32, 1)])
def test_warp_identity_shift_and_quarter_turn(dtype, c):
    rng = np.random.default_rng(5)
    a = _img(rng, 23, 31, c, dtype)
    out = np.zeros_like(a)
    assert np.array_equal(O.image_warp(a, I8, out), a)
    # h13, h23 = (3, 2): the inverse map takes output (x, y) to input (x - 3, y - 2)
    out = np.full_like(a, 7)
    O.image_warp(a, np.array([1, 0, 3, 0, 1, 2, 0, 0], np.float32), out)
    assert np.array_equal(out[2:, 3:], a[:-2, :-3])
    assert (out[:2] == 7).all() and (out[:, :3] == 7).all()   # untouched outside
    # forward (x, y) -> (W-1-y, x) on a square image: out(y, x) = in(W-1-x, y)
    s = a[:23, :23]
    sq = np.ascontiguousarray(s)
    out = np.zeros_like(sq)
    O.image_warp(sq, np.array([0, -1, 22, 1, 0, 0, 0, 0], np.float32), out)
    want = np.transpose(sq, (1, 0) + ((2,) if c > 1 else ()))[:, ::-1]
    assert np.array_equal(out, want)


def test_warp_truncates_bilinear_samples_into_uint8():
    a = np.array([[0, 100], [200, 255]], np.uint8)
    out = np.zeros((1, 1), np.uint8)
    O.image_warp(a, np.array([1, 0, -0.5, 0, 1, -0.5, 0, 0], np.float32), out)   # samples (0.5, 0.5)
    assert out[0, 0] == 138   # (0 + 100 + 200 + 255) / 4 = 138.75, truncated


def test_fuse_channels_and_homographies():
    rng = np.random.default_rng(9)
    ref, mov = _img(rng, 20, 24), _img(rng, 20, 24)
    fused = np.zeros((20, 24, 3), np.uint8)
    O.image_fuse(ref, mov, I8, I8, 0, 0, fused)
    assert np.array_equal(fused[..., 0], mov) and np.array_equal(fused[..., 1], ref)
    assert np.array_equal(fused[..., 2], mov)
    fused = np.zeros((20, 24, 3), np.uint8)
    O.image_fuse(ref, mov, np.array([1, 0, 1, 0, 1, 0, 0, 0], np.float32), I8, 0, 0, fused)
    assert np.array_equal(fused[:, 1:, 0], mov[:, :-1]) and (fused[:, 0, 0] == 0).all()
    assert np.array_equal(fused[..., 2], mov)


@pytest.mark.parametrize("size", [(16, 32), (64, 64), (8, 128)])
def test_fuse_local_identity_exact_on_power_of_two_sizes(size):
    h, w = size
    rng = np.random.default_rng(h * w)
    ref, mov = _img(rng, h, w), _img(rng, h, w)
    fused = np.zeros((h, w, 3), np.uint8)
    O.image_fuse(ref, mov, I8, I8, 0, 1, fused)
    # steps of 1/size are exact: step i sits at pixel i and samples it exactly
    assert np.array_equal(fused[..., 0], mov) and np.array_equal(fused[..., 1], ref)
    assert np.array_equal(fused[..., 2], mov)


def test_fuse_local_step_count_is_the_binary32_walk():
    # the normalised walk of a non-power-of-two axis, counted in numpy float32:
    # where it takes one step more than the size, the extra step's target is
    # past the last column and the row's pixel after it (the next row's first)
    # is written by it
    def walk(n):
        t, ts, s = np.float32(-0.5), [], np.float32(1) / np.float32(n)
        while t < np.float32(0.5):
            ts.append(t)
            t = np.float32(t + s)
        return ts
    # sizes whose walk takes a step more than the size, that step's point
    # (t + 0.5) * n (binary64, stored in binary32) still inside the image
    sizes = [n for n in range(2, 100)
             if len(walk(n)) == n + 1 and np.float32((float(walk(n)[-1]) + 0.5) * n) < n]
    assert sizes[:4] == [23, 24, 25, 29]
    n = sizes[0]
    rng = np.random.default_rng(n)
    ref = _img(rng, 8, n, dtype=np.float32) + np.float32(0.25)
    fused = np.zeros((8, n, 3), np.float32)
    O.image_fuse(ref, ref, I8, I8, 0, 1, fused)
    # the extra step samples x = (t + 0.5) * n just below n on row r: mostly the
    # next row's first pixel through the ceil tap's wrap, a little of ref[r, n-1]
    # -- and it lands on pixel (r + 1, 0), after that pixel's own step
    extra = fused[1:, 0, 1]
    assert not np.array_equal(extra, ref[1:, 0])
    assert np.allclose(extra, ref[1:, 0], atol=0.01 * 256)
    # column 0's own steps sample x = 0 exactly; every step lands within a
    # rounding error of its pixel
    assert fused[0, 0, 1] == ref[0, 0]
    assert np.allclose(fused[:, 1:, 1], ref[:, 1:], atol=0.05)


@pytest.mark.parametrize("c", [1, 2, 3, 4])
def test_write_png_round_trip(tmp_path, c):
    from PIL import Image
    from cugridsearch_amd import _lib as L
    rng = np.random.default_rng(c)
    a = rng.integers(0, 256, (300, 257, c) if c > 1 else (300, 257), dtype=np.uint8)   # > 64 KiB: several blocks
    p = tmp_path / f"x{c}.png"
    assert L.lib().cgs_write_png(str(p).encode(), a.ctypes.data, a.shape[1], a.shape[0], c) == 0
    with Image.open(p) as im:
        back = np.asarray(im)
    assert back.shape == a.shape and np.array_equal(back, a)
