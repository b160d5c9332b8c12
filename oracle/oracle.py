"""ctypes binding of the CPU oracle (cgs_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product (cugridsearch_amd)
never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# CGS_ORACLE_SO names another build of the same source (the sanitizer build
# of oracle/Makefile's `sanitize` target).
_SO = os.environ.get("CGS_ORACLE_SO") or os.path.join(_HERE, "_build", "libcgs_oracle.so")

COST_SAD, COST_NCC, COST_MI, COST_SQD, COST_SAD_SUM, COST_MI_PLAIN = range(6)
COST_NCC_STREAM, COST_SQD_STREAM, COST_SAD_STREAM = 6, 7, 8
COST_SQD_ALT, COST_NCC_ALT, COST_NCC_STREAM_ALT, COST_MI_STREAM_ALT, COST_MI_ALT = 9, 10, 11, 12, 13
ALT_COSTS = (COST_SQD_ALT, COST_NCC_ALT, COST_NCC_STREAM_ALT, COST_MI_STREAM_ALT, COST_MI_ALT)
NEAREST, BILINEAR = 0, 1
PARAM_CENTRED, PARAM_NONCENTRED, PARAM_RAW_H, PARAM_TRANSLATION, PARAM_NORM = range(5)
TIE_LOWEST, TIE_REFERENCE = 0, 1
PIX_U8, PIX_F32 = 0, 1


class _Image(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("width", ctypes.c_int), ("height", ctypes.c_int),
                ("channels", ctypes.c_int), ("pixtype", ctypes.c_int)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = ctypes.CDLL(_SO)
        fp = ctypes.POINTER(ctypes.c_float)
        i64 = ctypes.c_int64
        L.orc_grid_update.argtypes = [ctypes.c_int, fp, fp, fp, fp]
        L.orc_grid_num_elements.argtypes = [ctypes.c_int, fp]
        L.orc_grid_num_elements.restype = i64
        L.orc_grid_point.argtypes = [ctypes.c_int, fp, fp, fp, i64, fp]
        L.orc_homography.argtypes = [ctypes.c_int, fp, ctypes.c_float, ctypes.c_float, fp]
        L.orc_warp_coeffs.argtypes = [ctypes.c_int, fp, ctypes.c_float, ctypes.c_float, fp]
        L.orc_warp_point.argtypes = [fp, ctypes.c_int, ctypes.c_int, fp, ctypes.c_float, ctypes.c_float]
        L.orc_warp_point.restype = ctypes.c_int
        L.orc_sample_nearest.argtypes = [ctypes.POINTER(_Image), ctypes.c_float, ctypes.c_float, ctypes.c_int]
        L.orc_sample_nearest.restype = ctypes.c_float
        L.orc_sample_bilinear.argtypes = [ctypes.POINTER(_Image), ctypes.c_float, ctypes.c_float, ctypes.c_int]
        L.orc_sample_bilinear.restype = ctypes.c_float
        L.orc_cost_point.argtypes = [ctypes.POINTER(_Image), ctypes.POINTER(_Image), ctypes.c_int,
                                     ctypes.c_int, ctypes.c_int, fp, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_double)]
        L.orc_cost_point.restype = ctypes.c_float
        L.orc_search.argtypes = [ctypes.POINTER(_Image), ctypes.POINTER(_Image), ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int, ctypes.c_int, fp, fp, fp, ctypes.c_int, i64, i64, fp,
                                 ctypes.c_int]
        L.orc_reference_rank.argtypes = [i64]
        L.orc_reference_rank.restype = i64
        L.orc_argmin.argtypes = [fp, i64, ctypes.c_int, fp, ctypes.POINTER(i64)]
        L.orc_argmin_reference_sim.argtypes = [fp, i64, fp, ctypes.POINTER(i64)]
        L.orc_max_threads.restype = ctypes.c_int
        L.orc_covariance.argtypes = [ctypes.c_int, fp, fp, fp, fp, i64, ctypes.POINTER(ctypes.c_double), fp]
        L.orc_covariance.restype = i64
        L.orc_alt_homography.argtypes = [ctypes.c_int, fp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, fp]
        L.orc_alt_line.argtypes = [ctypes.c_int, fp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_float, fp]
        L.orc_mi_alt_rows.argtypes = [ctypes.c_int, fp, i64]
        L.orc_mi_alt_rows.restype = i64
        L.orc_mi_alt_area.argtypes = [ctypes.c_int, ctypes.c_int, i64]
        L.orc_mi_alt_area.restype = ctypes.c_float
        L.orc_image_filter.argtypes = [ctypes.POINTER(_Image), ctypes.c_void_p, fp, ctypes.c_int, ctypes.c_int]
        L.orc_image_quantize.argtypes = [ctypes.POINTER(ctypes.c_uint8), i64, ctypes.c_float]
        L.orc_image_resize.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_int]
        L.orc_image_resize.restype = ctypes.c_int
        L.orc_image_warp.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)]
        for nm in ("orc_matrix_multiply_f32", "orc_matrix_multiply_f64"):
            getattr(L, nm).argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int] * 3
        L.orc_image_fuse.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_float),
                                     ctypes.POINTER(ctypes.c_float), ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
        _lib = L
    return _lib


def _f32p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def image(arr):
    """Wraps an HxW or HxWxC uint8/float32 numpy array (row-major, channels
    interleaved, as CudaImage stores it)."""
    arr = np.ascontiguousarray(arr)
    if arr.dtype == np.uint8:
        pt = PIX_U8
    elif arr.dtype == np.float32:
        pt = PIX_F32
    else:
        raise TypeError("oracle images are uint8 or float32")
    h, w = arr.shape[:2]
    c = arr.shape[2] if arr.ndim == 3 else 1
    im = _Image(arr.ctypes.data, w, h, c, pt)
    im._keep = arr
    return im


def grid_update(start, end, count):
    s = np.ascontiguousarray(start, np.float32)
    e = np.ascontiguousarray(end, np.float32)
    c = np.array(count, np.float32)
    r = np.zeros_like(c)
    lib().orc_grid_update(len(c), _f32p(s), _f32p(e), _f32p(c), _f32p(r))
    return c, r


def grid_num_elements(count):
    c = np.ascontiguousarray(count, np.float32)
    return int(lib().orc_grid_num_elements(len(c), _f32p(c)))


def grid_point(start, res, count, index):
    s = np.ascontiguousarray(start, np.float32)
    r = np.ascontiguousarray(res, np.float32)
    c = np.ascontiguousarray(count, np.float32)
    out = np.zeros(len(c), np.float32)
    lib().orc_grid_point(len(c), _f32p(s), _f32p(r), _f32p(c), int(index), _f32p(out))
    return out


def homography(params, cx, cy, mode=PARAM_CENTRED):
    p = np.zeros(8, np.float32)
    p[:len(params)] = params
    h = np.zeros(8, np.float32)
    lib().orc_homography(mode, _f32p(p), cx, cy, _f32p(h))
    return h


def warp_coeffs(params, cx, cy, mode=PARAM_CENTRED):
    p = np.zeros(8, np.float32)
    p[:len(params)] = params
    c = np.zeros(9, np.float32)
    lib().orc_warp_coeffs(mode, _f32p(p), cx, cy, _f32p(c))
    return c


def warp_point(coef, x, y, mw=1e30, mh=1e30):
    """Returns ((X, Y) rounded sample point, inImage of the exact quotient)."""
    c = np.ascontiguousarray(coef, np.float32)
    out = np.zeros(2, np.float32)
    inside = lib().orc_warp_point(_f32p(c), int(x), int(y), _f32p(out), float(mw), float(mh))
    return out, bool(inside)


def alt_homography(params, mw, mh, fw, fh, norm=False):
    """H of the line-traversal costs: parametersToHomography (centred on the
    moving image) or parametersToHomographyNorm (norm=True)."""
    p = np.zeros(8, np.float32)
    p[:len(params)] = params
    h = np.zeros(8, np.float32)
    lib().orc_alt_homography(int(norm), _f32p(p), mw, mh, fw, fh, _f32p(h))
    return h


def alt_line(h, mw, mh, fw, fh, row, norm=False):
    """One fixed row's line: dict(ls, le, v, p0, inside)."""
    hh = np.ascontiguousarray(h, np.float32)
    out = np.zeros(7, np.float32)
    lib().orc_alt_line(int(norm), _f32p(hh), mw, mh, fw, fh, float(row), _f32p(out))
    return dict(ls=out[0], le=out[1], v=out[2:4].copy(), p0=out[4:6].copy(), inside=bool(out[6]))


def mi_alt_rows(fh):
    n = int(lib().orc_mi_alt_rows(fh, None, 0))
    rows = np.zeros(n, np.float32)
    lib().orc_mi_alt_rows(fh, _f32p(rows), n)
    return rows


def mi_alt_area(fw, fh, n):
    return np.float32(lib().orc_mi_alt_area(fw, fh, int(n)))


def image_filter(img, filt):
    """filterProcess with every channel filtered; filt is a (fh, fw) array."""
    im = image(img)
    f = np.ascontiguousarray(filt, np.float32)
    out = np.zeros_like(np.ascontiguousarray(img))
    lib().orc_image_filter(ctypes.byref(im), out.ctypes.data, _f32p(f), f.shape[1], f.shape[0])
    return out


def image_quantize(img, mod):
    """transformPixel on a uint8 image (a copy is returned)."""
    out = np.array(img, np.uint8, copy=True, order="C")
    lib().orc_image_quantize(out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), out.size, float(mod))
    return out


def image_resize(img, ow, oh):
    """stbir_resize_uint8 of an (H, W) or (H, W, C) uint8 image to (oh, ow)."""
    a = np.ascontiguousarray(img, np.uint8)
    C = 1 if a.ndim == 2 else a.shape[2]
    out = np.zeros((oh, ow) + (() if a.ndim == 2 else (C,)), np.uint8)
    if lib().orc_image_resize(a.ctypes.data, a.shape[1], a.shape[0], out.ctypes.data, ow, oh, C) != 0:
        raise ValueError("bad resize sizes")
    return out


def image_warp(img, H, out):
    """transformImage of img into out (a numpy array it writes in place)."""
    im, om = image(img), image(out)
    lib().orc_image_warp(ctypes.byref(im), ctypes.byref(om), _f32p(np.ascontiguousarray(H, np.float32)))
    return out


def image_fuse(ref, mov, est, truth, channel, local, fused):
    """fuseAlignedImages(_local) into fused, an (H, W, 3) uint8 array written in place."""
    rm, mm, fm = image(ref), image(mov), image(fused)
    lib().orc_image_fuse(ctypes.byref(rm), ctypes.byref(mm), _f32p(np.ascontiguousarray(est, np.float32)),
                         _f32p(np.ascontiguousarray(truth, np.float32)), int(channel), int(local), ctypes.byref(fm))
    return fused


def matrix_multiply(A, B):
    """CudaMatrix::multiply of float32 or float64 row-major matrices."""
    A = np.ascontiguousarray(A)
    B = np.ascontiguousarray(B, A.dtype)
    C = np.zeros((A.shape[0], B.shape[1]), A.dtype)
    f = lib().orc_matrix_multiply_f32 if A.dtype == np.float32 else lib().orc_matrix_multiply_f64
    f(A.ctypes.data, B.ctypes.data, C.ctypes.data, A.shape[0], A.shape[1], B.shape[1])
    return C


def sample(img, y, x, sampler, c=0):
    im = image(img)
    f = lib().orc_sample_nearest if sampler == NEAREST else lib().orc_sample_bilinear
    return f(ctypes.byref(im), float(y), float(x), c)


def cost_point(moved, fixed, cost, sampler, params, mode=PARAM_CENTRED, bins=32, want_hist=False):
    """Returns (cost, aux) with aux = [sum, i1, i2, ic, count, mi, hx, hy, mi_cost]
    (and the bins x bins histogram when want_hist).  For the line-traversal
    costs aux = [steps, sum f^2, sum m^2, cross term, count, mi, moved area,
    fixed area, overlap test passed]."""
    m, f = image(moved), image(fixed)
    p = np.zeros(8, np.float32)
    p[:len(params)] = params
    aux = np.zeros(16, np.float64)
    hist = np.zeros(bins * bins, np.uint32) if want_hist else None
    hp = hist.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)) if want_hist else None
    v = lib().orc_cost_point(ctypes.byref(m), ctypes.byref(f), cost, sampler, mode, _f32p(p), bins, hp,
                             aux.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    if want_hist:
        return v, aux, hist.reshape(bins, bins)
    return v, aux


def search(moved, fixed, cost, sampler, start, res, count, mode=PARAM_CENTRED, bins=32,
           i0=0, i1=None, nthreads=0):
    m, f = image(moved), image(fixed)
    s = np.ascontiguousarray(start, np.float32)
    r = np.ascontiguousarray(res, np.float32)
    c = np.ascontiguousarray(count, np.float32)
    n = grid_num_elements(c)
    if i1 is None:
        i1 = n
    out = np.zeros(i1 - i0, np.float32)
    lib().orc_search(ctypes.byref(m), ctypes.byref(f), cost, sampler, mode, len(c), _f32p(s), _f32p(r),
                     _f32p(c), bins, i0, i1, _f32p(out), nthreads)
    return out


def reference_rank(idx):
    return int(lib().orc_reference_rank(int(idx)))


def argmin(r, tie_mode=TIE_LOWEST):
    a = np.ascontiguousarray(r, np.float32)
    v = ctypes.c_float()
    i = ctypes.c_int64()
    lib().orc_argmin(_f32p(a), len(a), tie_mode, ctypes.byref(v), ctypes.byref(i))
    return v.value, i.value


def argmin_reference_sim(r):
    a = np.ascontiguousarray(r, np.float32)
    v = ctypes.c_float()
    i = ctypes.c_int64()
    lib().orc_argmin_reference_sim(_f32p(a), len(a), ctypes.byref(v), ctypes.byref(i))
    return v.value, i.value


def covariance(start, res, count, result):
    """calculateCovariance_by_value restated: (D x D covariance, min index, min value)."""
    s = np.ascontiguousarray(start, np.float32)
    r = np.ascontiguousarray(res, np.float32)
    c = np.ascontiguousarray(count, np.float32)
    v = np.ascontiguousarray(result, np.float32)
    D = len(c)
    out = np.zeros(D * D, np.float64)
    mv = ctypes.c_float()
    mi = lib().orc_covariance(D, _f32p(s), _f32p(r), _f32p(c), _f32p(v), len(v),
                              out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(mv))
    return out.reshape(D, D), int(mi), mv.value


def max_threads():
    return int(lib().orc_max_threads())


# ------------------------------------------------------- proposal network

_CNN_SO = os.path.join(_HERE, "_build", "libcnn_oracle.so")
_cnn = None


def cnn_lib():
    global _cnn
    if _cnn is None:
        if not os.path.exists(_CNN_SO):
            build()
        L = ctypes.CDLL(_CNN_SO)
        dp = ctypes.POINTER(ctypes.c_double)
        L.cnn_param_count.restype = ctypes.c_int64
        L.cnn_round_bf16.argtypes = [ctypes.c_double]
        L.cnn_round_bf16.restype = ctypes.c_double
        L.cnn_loss_grad.argtypes = [dp, dp, dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp]
        L.cnn_loss_grad.restype = ctypes.c_double
        _cnn = L
    return _cnn


def cnn_param_count():
    return int(cnn_lib().cnn_param_count())


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def cnn_loss_grad(params, x, target=None, emulate=True, want_grad=True):
    """Network forward (and backward) in binary64.  x: [n, S, S, 2];
    target: [n, 8] or None.  emulate: round weights, activations and
    inter-layer gradients to bf16 where the GPU stores bf16.
    Returns (loss, out [n, 8], grad [P] or None)."""
    p = np.ascontiguousarray(params, np.float64)
    xx = np.ascontiguousarray(x, np.float64)
    n, side = xx.shape[0], xx.shape[1]
    out = np.zeros((n, 8), np.float64)
    t = None if target is None else np.ascontiguousarray(target, np.float64)
    g = np.zeros(cnn_param_count(), np.float64) if (want_grad and t is not None) else None
    loss = cnn_lib().cnn_loss_grad(_dp(p), _dp(xx), _dp(t) if t is not None else None, n, side, 1 if emulate else 0,
                                   _dp(out), _dp(g) if g is not None else None)
    return loss, out, g


def round_bf16(a):
    """binary64 -> binary32 -> bf16 (nearest even), as float64."""
    f = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    nan = (f & 0x7FFFFFFF) > 0x7F800000
    r = ((f + 0x7FFF + ((f >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    out = r.view(np.float32).astype(np.float64)
    return np.where(nan, np.asarray(a, np.float64), out)
