"""CPU checks of the multispectral matcher's pieces: the oracle's image filter
and quantisation against independent numpy restatements and hand-computed
answers, and the matcher's grid set-up, window update and homography
conversions against the reference's formulas worked by hand."""
import math

import numpy as np
import pytest

from cugridsearch_amd import matcher as M


def _filter_numpy(img, filt):
    """float64 restatement of filterProcess: outside taps read the centre
    pixel, the running sum carries over channels, uint8 output saturated."""
    h, w = img.shape[:2]
    x = img.reshape(h, w, -1).astype(np.float64)
    C = x.shape[2]
    fh, fw = filt.shape
    out = np.zeros_like(x)
    for yy in range(h):
        for xx in range(w):
            acc = 0.0
            for c in range(C):
                for r in range(fh):
                    for k in range(fw):
                        ty, tx = yy - (fh - 1) // 2 + r, xx - (fw - 1) // 2 + k
                        v = x[ty, tx, c] if (0 <= ty < h and 0 <= tx < w) else x[yy, xx, c]
                        acc += float(filt[r, k]) * v
                out[yy, xx, c] = acc
    if img.dtype == np.uint8:
        out = np.clip(np.trunc(out), 0, 255)
    return out.reshape(img.shape)


def test_filter_unsharp_brute_force(oracle):
    rng = np.random.default_rng(1)
    img = rng.integers(0, 256, (13, 17)).astype(np.uint8)
    got = oracle.image_filter(img, M.UNSHARP_3X3)
    assert np.array_equal(got, _filter_numpy(img, M.UNSHARP_3X3).astype(np.uint8))   # integers: exact
    # a constant image stays constant (5 - 4 = 1, outside taps read the centre)
    flat = np.full((6, 9), 77, np.uint8)
    assert np.array_equal(oracle.image_filter(flat, M.UNSHARP_3X3), flat)
    # an isolated bright pixel: 5 * 255 saturates to 255, its neighbours -255 to 0
    spot = np.zeros((5, 5), np.uint8)
    spot[2, 2] = 255
    out = oracle.image_filter(spot, M.UNSHARP_3X3)
    assert out[2, 2] == 255 and out[1, 2] == 0 and out[2, 1] == 0 and out.sum() == 255
    # 1x1 delta: the identity
    assert np.array_equal(oracle.image_filter(img, np.ones((1, 1), np.float32)), img)


def test_filter_even_sizes_channels_and_floats(oracle):
    rng = np.random.default_rng(2)
    img = rng.random((11, 14)).astype(np.float32) * 255
    for shape in ((5, 5), (10, 10), (3, 4)):
        filt = rng.random(shape).astype(np.float32) - np.float32(0.3)
        got = oracle.image_filter(img, filt)
        want = _filter_numpy(img, filt)
        assert np.allclose(got, want, rtol=1e-5, atol=1e-3), shape
    # two channels: the second channel's output also holds the first's sum
    two = rng.integers(0, 40, (7, 8, 2)).astype(np.uint8)
    got = oracle.image_filter(two, M.UNSHARP_3X3)
    assert np.array_equal(got, _filter_numpy(two, M.UNSHARP_3X3).astype(np.uint8))
    ch0 = _filter_numpy(two[..., :1].copy(), M.UNSHARP_3X3)
    assert np.array_equal(got[..., 0], ch0[..., 0].astype(np.uint8))


def test_quantize_known_answers(oracle):
    v = np.arange(256, dtype=np.uint8)
    assert np.array_equal(oracle.image_quantize(v, 1.0), v)          # through a signed char and back
    q = oracle.image_quantize(v, 10.0)
    assert q[37] == 30 and q[127] == 120 and q[128] == 136 and q[200] == 206 and q[255] == 0
    # 200 -> (char) -56 -> -5.6 -> -5 -> -50 -> 206; 128 -> -128 -> -12 -> -120 -> 136
    q = oracle.image_quantize(v, 2.5)
    assert q[7] == 5 and q[8] == 7   # 7 / 2.5 = 2.8 -> 2 -> 5.0; 8 / 2.5 = 3.2 -> 3 -> 7.5 -> 7
    sv = v.astype(np.int8).astype(np.float32)
    for mod in (0.3, 3.0, 7.5, 100.0):
        t = np.trunc(sv / np.float32(mod)).astype(np.float32)
        want = (np.trunc(t * np.float32(mod)).astype(np.int64) & 255).astype(np.uint8)
        assert np.array_equal(oracle.image_quantize(v, mod), want), mod


def test_grid_set_up_from_the_reference_text():
    num, start, end = M.coarse_grid(256, 256, 256, 256)
    assert [float(x) for x in num[:4]] == [32, 16, 16, 5]
    assert num[4] == np.float32(257) / np.float32(20) and num[4] == np.float32(12.85)
    assert start[1] == np.float32(1.5) and end[1] == np.float32(5) and start[3] == np.float32(-0.4)
    assert start[4] == -128 and end[4] == 128 and end[5] == 128
    assert end[0] == np.float32(2 * 3.14159265 - 3.14159265 / 32)
    # the window update reads the new start when it sets the end
    s, e = M.shrink_window([0.0] * 8, [8.0] * 8, [np.float32(4)] * 8)
    assert s[0] == 2.0 and e[0] == np.float32(4 + (8 - 2) / 4)
    cnum, cs, ce = M.covariance_grid([np.float32(v) for v in (0.5, 2, 3, 0.1, 10, -4, 0, 0)])
    assert cs[0] == np.float32(0.5 - 3.14159265 / 64) and ce[4] == np.float32(12.5) and cs[5] == np.float32(-6.5)
    assert cs[1] == ce[1] == np.float32(2) and [float(x) for x in cnum] == [5, 1, 1, 1, 5, 5, 1, 1]


def test_homography_round_trip():
    """parametersToHomographyNorm then homographyToParameters: with equal
    image sizes the rotation, scales and shear come back, and the
    translations come back as h13 / h23 (the reference's conversion)."""
    for p in ([0.3, 1.2, 0.9, 0.1, 5.0, -3.0, 1e-3, -2e-3], [-1.2, 2.0, 1.5, -0.2, 0.0, 0.0, 0.0, 0.0]):
        p = [np.float32(v) for v in p]
        h = M.parameters_to_homography_norm(p, 64, 48, 64, 48)
        back = M.homography_to_parameters(h)
        assert abs(back[0] - p[0]) < 1e-5 and abs(back[1] - p[1]) < 1e-5
        assert abs(back[2] - p[2]) < 1e-4 and abs(back[3] - p[3]) < 1e-4
        assert back[4] == h[2] and back[5] == h[5] and back[6] == p[6] and back[7] == p[7]
    # the integer size ratio: a fixed image twice as wide halves scaleX
    h = M.parameters_to_homography_norm([np.float32(0), np.float32(1), np.float32(1)] + [np.float32(0)] * 5,
                                        32, 32, 64, 32)
    assert h[0] == np.float32(0.5) and h[4] == np.float32(1)


def test_protocol_formatting():
    assert M.g(np.float32(1.5)) == "1.5" and M.g(np.float32(1e-7)) == "1e-07" and M.g(123456789.0) == "1.23457e+08"
    assert M.eigen_column([1.5, -0.25, 10]) == "  1.5\n-0.25\n   10"
    banner = M.sobel_banner().splitlines()
    assert len(banner) == 25 and banner[2] == "1, 0, 0, 0, 0, "


def test_nan_renders_with_its_sign():
    """glibc prints a NaN whose sign bit is set as "-nan" (what the
    reference's % Error column shows for 0/0 on x86), and others as "nan"."""
    z = np.float32(0)
    with np.errstate(invalid="ignore"):
        q = z / z
    assert M.g(q) == ("-nan" if math.copysign(1.0, float(q)) < 0 else "nan")
    assert M.g(float("nan")) == "nan" and M.g(-float("nan")) == "-nan"
    assert M.g(float("inf")) == "inf" and M.g(-float("inf")) == "-inf"


def test_image_number_from_the_reference_rule():
    """src/cpu/multispectral_ImageMatcher.cu:991-999: the characters after the
    last one outside "0123456789.png", up to the next '.', through atoi.
    The set holds 'p', 'n' and 'g', so a name whose digits follow one of those
    letters reads 0 ("img12.png" -> "g12"), as does a ".jpg" name ("pg")."""
    cases = {"fixed_0007.png": 7, "/data/set_2/im_12.png": 12, "/data/set_2/img12.png": 0, "png_3.png": 3,
             "x/7.png": 7, "7.png": 7, "a_ng5.png": 0, "scan-0042.jpg": 0, "name_.png": 0, "noext_15": 15,
             "data/img_2_0009.png": 9}
    for name, want in cases.items():
        assert M.image_number(name) == want, name
    assert M.atoi("  -12abc") == -12 and M.atoi("+7") == 7 and M.atoi("x1") == 0 and M.atoi("") == 0


def test_stof_is_strtof():
    """std::stof: one rounding, decimal straight to binary32 (not through
    binary64), leading white space skipped, trailing text ignored, and an
    error for no number or a result out of range."""
    assert M.stof("1.5") == np.float32(1.5) and M.stof("  -2.25,3") == np.float32(-2.25)
    assert M.stof("1e3x") == np.float32(1000) and M.stof("0x1p-3") == np.float32(0.125)
    # a decimal between two binary32 values whose binary64 rounding lands on
    # their midpoint: rounding twice goes the wrong way, strtof does not
    tie = "1.000000059604644775390625000001"     # just above 1 + 2^-24
    assert float(np.float32(float(tie))) == 1.0
    assert M.stof(tie) == np.float32(1.0000001192092896)
    for bad in ("", ",1", "abc", "1e-50", "1e50"):
        with pytest.raises(ValueError):
            M.stof(bad)


def test_ground_truth_lookup(tmp_path):
    """src/cpu/multispectral_ImageMatcher.cu:988-1031: the matching row's
    eight values, the scale factors on the first six, the wrap past a short
    row's end, and no scaling when the file does not open."""
    H = [np.float32(v) for v in (1, 2, 3, 4, 5, 6, 7, 8)]
    sx, sy = np.float32(0.5), np.float32(0.25)
    p = tmp_path / "gt.csv"
    p.write_text("id,h11,h12\n3,9,9,9,9,9,9,9,9\n7,1.5,0.1,-3,4e-2,5,6,0.001,-0.002\n7,0,0,0,0,0,0,0,0\n")
    got = M.ground_truth(str(p), "dir/fixed_0007.png", sx, sy, H)
    vals = [np.float32(v) for v in ("1.5", "0.1", "-3", "4e-2", "5", "6", "0.001", "-0.002")]
    want = [np.float32(v * sx) for v in vals[:3]] + [np.float32(v * sy) for v in vals[3:6]] + vals[6:]
    assert [np.float32(v) for v in got] == want
    # a short row wraps to its own start: "7,1.5,2" reads 1.5, 2, 7, 1.5, 2, 7, ...
    p.write_text("7,1.5,2\n")
    got = M.ground_truth(str(p), "fixed_7.png", np.float32(1), np.float32(1), H)
    assert [float(v) for v in got] == [1.5, 2, 7, 1.5, 2, 7, 1.5, 2]
    # no row matches: the estimate, scaled; no file: the estimate as it is
    p.write_text("3,1,1,1,1,1,1,1,1\n")
    got = M.ground_truth(str(p), "fixed_0007.png", sx, sy, H)
    assert [np.float32(v) for v in got] == [np.float32(v * sx) for v in H[:3]] + \
        [np.float32(v * sy) for v in H[3:6]] + H[6:]
    assert M.ground_truth(str(tmp_path / "none.csv"), "fixed_0007.png", sx, sy, H) == H
    # a field that is no number: std::stof throws in the reference
    p.write_text("7,1,2,x,4,5,6,7,8\n")
    with pytest.raises(ValueError):
        M.ground_truth(str(p), "fixed_0007.png", sx, sy, H)
