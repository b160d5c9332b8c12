"""The resize restatement (oracle/cgs_oracle.c, orc_image_resize) against the
library the reference calls: stb_image_resize v0.97 built from the
reference's own header (include/third_party/stb_image_resize.h) into
oracle/_ref by `make -C oracle ref`.  Bit for bit on every case: random sizes
both ways on each axis, one to four channels, extreme ratios, the identity
size (not a copy: the Mitchell filter still blurs) and the matcher's shapes."""
import ctypes
import os

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(HERE, "..", "oracle", "_ref", "libstbir_ref.so")

pytestmark = pytest.mark.skipif(not os.path.exists(REF),
                                reason="oracle/_ref not built (needs the reference tree; make -C oracle ref)")


def _ref():
    lib = ctypes.CDLL(REF)
    lib.stbir_resize_uint8.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.stbir_resize_uint8.restype = ctypes.c_int
    return lib


def ref_resize(a, ow, oh):
    a = np.ascontiguousarray(a, np.uint8)
    C = 1 if a.ndim == 2 else a.shape[2]
    out = np.zeros((oh, ow) + (() if a.ndim == 2 else (C,)), np.uint8)
    assert _ref().stbir_resize_uint8(a.ctypes.data, a.shape[1], a.shape[0], 0, out.ctypes.data, ow, oh, 0, C) == 1
    return out


def _image(rng, h, w, c, smooth):
    a = rng.integers(0, 256, (h, w, c) if c > 1 else (h, w), dtype=np.uint8)
    if smooth:   # ramps exercise the interpolation, noise the ringing and clamping
        a = (np.cumsum(a.astype(np.int64), axis=1) // max(w, 1) % 256).astype(np.uint8)
    return a


def test_random_sizes_bit_exact():
    rng = np.random.default_rng(2024)
    for _ in range(300):
        iw, ih, ow, oh = (int(v) for v in rng.integers(1, 90, 4))
        c = int(rng.integers(1, 5))
        a = _image(rng, ih, iw, c, rng.random() < 0.5)
        assert np.array_equal(O.image_resize(a, ow, oh), ref_resize(a, ow, oh)), (iw, ih, ow, oh, c)


@pytest.mark.parametrize("shape", [((700, 3), (2, 5)), ((3, 700), (5, 2)), ((1, 1), (64, 33)), ((64, 33), (1, 1)),
                                   ((97, 61), (97, 61)), ((512, 384), (256, 256)), ((256, 256), (512, 384)),
                                   ((640, 480), (320, 200)), ((300, 100), (100, 200))],
                         ids=lambda s: f"{s[0][0]}x{s[0][1]}->{s[1][0]}x{s[1][1]}")
def test_extreme_and_matcher_shapes(shape):
    (iw, ih), (ow, oh) = shape
    rng = np.random.default_rng(iw * 1000 + ih)
    for smooth in (False, True):
        a = _image(rng, ih, iw, 1, smooth)
        assert np.array_equal(O.image_resize(a, ow, oh), ref_resize(a, ow, oh))


def test_constant_image_stays_constant():
    # normalised weights: a flat image resamples to itself at any size
    for v in (0, 1, 128, 254, 255):
        a = np.full((37, 53), v, np.uint8)
        for ow, oh in ((20, 11), (53, 37), (100, 80)):
            got = O.image_resize(a, ow, oh)
            assert np.array_equal(got, ref_resize(a, ow, oh))
            assert (got == v).all(), (v, ow, oh)
