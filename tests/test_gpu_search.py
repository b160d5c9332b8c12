"""GPU parity of the plain search against the CPU oracle, through the C-ABI.

Bars: bit-exact costs and argmin for integer nearest-neighbour SAD; MI joint
histogram counts bit-exact; float costs (bilinear SAD / NCC / SQD, MI
entropy) within the relative tolerances written in each test."""
import json
import math
import os

import numpy as np
import pytest

from cugridsearch_amd import synth
from cugridsearch_amd import api as A
from cugridsearch_amd import _lib as L

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "workloads.json")
NCC_RTOL = 1e-5       # bilinear FP32 costs vs the FP64-accumulating oracle
MI_RTOL = 1e-6        # MI cost: FP64 entropy on both sides, rounded to FP32


def gpu_search(functor, moved, fixed, start, end, count, tie=L.TIE_LOWEST):
    grid = A.CudaGrid(len(count))
    grid.setStartPoint(start)
    grid.setEndPoint(end)
    grid.setNumSamples(count)
    res = A.CudaTensor(grid.getAxisSampleCounts())
    searcher = A.CudaGridSearcher(grid, res)
    v, i = searcher.search_by_value(functor, A.CudaImage.from_numpy(moved), A.CudaImage.from_numpy(fixed),
                                    tie_mode=tie)
    return res.data.cpu().numpy(), (v, i), grid


def oracle_search(oracle, functor, moved, fixed, start, end, count):
    c2, r = oracle.grid_update(start, end, count)
    return oracle.search(moved, fixed, functor.cost, functor.sampler, start, r, c2, mode=functor.param_mode,
                         bins=functor.bins)


def close_mask(a, b, rtol, atol=0.0):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    both_nan = np.isnan(a) & np.isnan(b)
    with np.errstate(invalid="ignore"):
        return both_nan | (a == b) | (np.abs(a - b) <= rtol * np.abs(b) + atol)


def close(a, b, rtol, atol=0.0):
    ok = close_mask(a, b, rtol, atol)
    return bool(ok.all()), np.where(~ok)[0][:10]


def test_nn_sad_64_bit_exact(gpu, oracle):
    m, f, s, e, c, spec = synth.workload("nn_sad_64")
    fn = A.centredAbsoluteDifference.with_(sampler=L.SAMPLER_NEAREST)
    got, (v, i), _ = gpu_search(fn, m, f, s, e, c)
    want = oracle_search(oracle, fn, m, f, s, e, c)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert (v, i) == oracle.argmin(want)
    got_r, (vr, ir), _ = gpu_search(fn, m, f, s, e, c, tie=L.TIE_REFERENCE)
    assert (vr, ir) == oracle.argmin(want, oracle.TIE_REFERENCE)


def test_nn_sad_ties_reference_order(gpu, oracle):
    """Identical flat images: every in-range point costs 0, so the argmin is
    decided by the tie order alone."""
    img = np.full((16, 16), 9, np.uint8)
    s, e, c = [-2, -2], [2, 2], [40, 40]
    fn = A.sumOfAbsoluteDifferences.with_(sampler=L.SAMPLER_NEAREST)
    got, (v, i), _ = gpu_search(fn, img, img, s, e, c, tie=L.TIE_REFERENCE)
    want = oracle_search(oracle, fn, img, img, s, e, c)
    assert np.array_equal(got, want)
    assert (v, i) == oracle.argmin(want, oracle.TIE_REFERENCE) == oracle.argmin_reference_sim(want)
    _, (v2, i2), _ = gpu_search(fn, img, img, s, e, c, tie=L.TIE_LOWEST)
    assert (v2, i2) == oracle.argmin(want)


def test_bilinear_ncc_64_full(gpu, oracle):
    m, f, _ = synth.make_pair(21, 64, sigma=3.0, perturb=3.0, dtype=np.float32)
    s, e, c = synth.grid_ranges(4, 4.0)
    got, (v, i), _ = gpu_search(A.calcNCC, m, f, s, e, c)
    want = oracle_search(oracle, A.calcNCC, m, f, s, e, c)
    ok, bad = close(got, want, NCC_RTOL)
    assert ok, (bad, got[bad], want[bad])
    vo, io = oracle.argmin(want)
    assert abs(v - vo) <= NCC_RTOL * abs(vo)
    assert abs(want[i] - vo) <= 2 * NCC_RTOL * abs(vo)   # same argmin up to near-ties


@pytest.mark.parametrize("name", ["bilinear_ncc_256", "bilinear_mi32_512"])
def test_full_workload_sample_and_golden_argmin(gpu, oracle, name):
    m, f, s, e, c, spec = synth.workload(name)
    fn = A.calcNCC if spec["cost"] == "ncc" else A.calcMIstream.with_(bins=spec["bins"])
    got, (v, i), grid = gpu_search(fn, m, f, s, e, c)
    n = grid.numElements()
    assert got.shape == (n,)
    rng = np.random.default_rng(5)
    idx = rng.choice(n, 64, replace=False)
    rtol = NCC_RTOL if spec["cost"] == "ncc" else MI_RTOL
    for k in idx:
        want, _ = oracle.cost_point(m, f, fn.cost, fn.sampler, grid.getGridPoint(k), bins=fn.bins)
        ok, _ = close([got[k]], [want], rtol)
        assert ok, (k, got[k], want)
    gold = json.load(open(GOLDEN))[name]
    assert gold["num_elements"] == n
    vo = gold["argmin_value"]
    assert abs(v - vo) <= rtol * abs(vo), (v, vo)
    assert abs(got[gold["argmin_index"]] - vo) <= rtol * abs(vo)
    ov, _ = oracle.cost_point(m, f, fn.cost, fn.sampler, grid.getGridPoint(i), bins=fn.bins)
    assert abs(ov - vo) <= 2 * rtol * abs(vo)
    # every committed full-grid oracle sample (2048 seeded indices)
    si = np.array(gold["sample_index"])
    ok, bad = close(got[si], np.array(gold["sample_value"], np.float64), rtol)
    assert ok, (si[bad], got[si[bad]], np.array(gold["sample_value"])[bad])
    # the oracle's 16 lowest costs: the GPU's value at each, and the GPU's own
    # 16 lowest values equal them (the same set up to near-ties)
    bi = np.array(gold["best16_index"])
    bv = np.array(gold["best16_value"], np.float64)
    ok, bad = close(got[bi], bv, rtol)
    assert ok, (bi[bad], got[bi[bad]], bv[bad])
    mine = np.sort(got[np.isfinite(got)])[:16].astype(np.float64)
    ok, bad = close(mine, bv, 2 * rtol)
    assert ok, (mine[bad], bv[bad])
    assert int(np.isnan(got).sum()) == gold["nan_count"]


def test_mi32_512_histograms_bit_exact(gpu, oracle):
    """The C3 workload (512x512 uint8, 32 bins): joint histograms of 8 grid
    points, the golden argmin among them, bit for bit against the oracle."""
    m, f, s, e, c, spec = synth.workload("bilinear_mi32_512")
    gold = json.load(open(GOLDEN))["bilinear_mi32_512"]
    grid = A.CudaGrid(8)
    grid.setStartPoint(s)
    grid.setEndPoint(e)
    grid.setNumSamples(c)
    gm, gf = A.CudaImage.from_numpy(m), A.CudaImage.from_numpy(f)
    fn = A.calcMIstream.with_(bins=spec["bins"])
    rng = np.random.default_rng(11)
    for k in [gold["argmin_index"]] + [int(x) for x in rng.choice(grid.numElements(), 7, replace=False)]:
        p = grid.getGridPoint(k)
        cost, hist = A.eval_point(fn, gm, gf, p)
        oc, _, oh = oracle.cost_point(m, f, fn.cost, fn.sampler, p, bins=fn.bins, want_hist=True)
        assert hist.sum() > 0
        assert np.array_equal(hist, oh), k
        assert close([cost], [oc], MI_RTOL)[0], (k, cost, oc)


def test_mi_in_image_on_rounded_quotient(gpu, oracle):
    """The 2^-24 window case at the moved image's lower edge (Y/W just below
    48 that rounds to 48.0): the MI costs leave it out (inImage on the
    rounded quotient), SAD/NCC/SQD keep it (exact quotient); the tiled search
    and the single-point path both agree with the oracle, with W in
    [1/64, 64] over the fixed image (h32 = 1: the refined-reciprocal MI
    sweep), with W up to 101 (h32 = 100: the IEEE-division sweep), and with a
    nonzero X coefficient below 2^-60 (h12 = 1e-25: the IEEE-division sweep
    although W is in range)."""
    rng = np.random.default_rng(8)
    moved = rng.integers(0, 256, (48, 64), dtype=np.uint8)
    fixed = rng.integers(0, 256, (2, 3), dtype=np.uint8)
    gm, gf = A.CudaImage.from_numpy(moved), A.CudaImage.from_numpy(fixed)
    for h32, h12 in ((1.0, 0.0), (100.0, 0.0), (1.0, 1e-25)):
        p = np.array([0, h12, 10, 0, 0, 48 + 2 ** -18, 2 ** -23, h32], np.float32)
        for base, want_count in ((A.calcMIstream, 4), (A.calcMI, 4), (A.calcNCC, 5), (A.calcSQD, 5),
                                 (A.averageAbsoluteDifference, 5)):
            fn = base.with_(param_mode=L.PARAM_RAW_H, bins=32)
            oc, aux = oracle.cost_point(moved, fixed, fn.cost, fn.sampler, p, mode=L.PARAM_RAW_H, bins=32)[:2]
            assert aux[4] == want_count, (h32, fn.name, aux[4])
            cost, hist = A.eval_point(fn, gm, gf, p)
            if hist is not None:
                assert hist.sum() == want_count, (h32, fn.name, hist.sum())
            got, (v, i), _ = gpu_search(fn, moved, fixed, p, p, [1] * 8)
            # MI: exact quotients; SAD/NCC/SQD (both calls run the tiled
            # kernel): the sample point is Xn * v_rcp(W), an ulp or so off the
            # rounded quotient, which on 5 samples moves the cost by up to
            # ~1e-4 (a wrong count moves it by 20% or more)
            rtol = MI_RTOL if hist is not None else 1e-3
            for where, c in (("eval_point", cost), ("search", got[0])):
                assert close([c], [oc], rtol)[0], (h32, fn.name, where, c, oc)


def _search_with(functor, m, f, grid, gpt, order):
    """cgs_search with an explicit tile size and work order (CGS_WORK_ORDER
    is read by every search call); returns the costs and the argmin."""
    import ctypes
    old = os.environ.get("CGS_WORK_ORDER")
    if order is None:
        os.environ.pop("CGS_WORK_ORDER", None)
    else:
        os.environ["CGS_WORK_ORDER"] = order
    try:
        o = A.make_opts(functor)
        o.tile_gpt = gpt
        out = L.Argmin()
        res = A.CudaTensor(grid.getAxisSampleCounts())
        L.check(L.lib().cgs_search(ctypes.byref(grid._g), ctypes.byref(m._img), ctypes.byref(f._img), ctypes.byref(o),
                                   ctypes.c_void_p(res.data.data_ptr()), ctypes.byref(out), None))
        return res.data.cpu().numpy(), (out.value, out.index)
    finally:
        if old is None:
            os.environ.pop("CGS_WORK_ORDER", None)
        else:
            os.environ["CGS_WORK_ORDER"] = old


@pytest.mark.parametrize("counts", [(3, 2, 2, 2, 6, 2, 2, 2), (3, 1, 1, 1, 5, 1, 1, 3)], ids=["box", "ragged45"])
def test_work_order_and_tiles_bit_identical(gpu, oracle, counts):
    """Which grid points share a workgroup (the work order, tx-first by
    default for the centred parametrisation) and how many (tile_gpt) change
    no cost bit and no argmin: the grid order with every tile size, the
    default order and a reversed order all agree, and match the oracle."""
    m, f, _ = synth.make_pair(23, 64, sigma=3.0, perturb=3.0, dtype=np.float32)
    s, e, _ = synth.grid_ranges(4, 4.0)
    grid = A.CudaGrid(8)
    grid.setStartPoint(s)
    grid.setEndPoint(e)
    grid.setNumSamples(np.array(counts, np.float32))
    mi, fi = A.CudaImage.from_numpy(m), A.CudaImage.from_numpy(f)
    for fn in (A.calcNCC, A.calcSQD, A.centredAbsoluteDifference, A.calcMIstream.with_(bins=32)):
        want, wam = _search_with(fn, mi, fi, grid, 4, "grid")
        for gpt, order in [(0, None), (6, None), (8, None), (6, "grid"), (8, "grid"), (4, "7,6,5,4,3,2,1,0"),
                           (2, None)]:
            got, am = _search_with(fn, mi, fi, grid, gpt, order)
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (fn.name, gpt, order)
            assert am == wam, (fn.name, gpt, order)
        ref = oracle_search(oracle, fn, m, f, s, e, np.array(counts, np.float32))
        ok, bad = close(want, ref, MI_RTOL if fn.cost == L.COST_MI else NCC_RTOL)
        assert ok, (fn.name, bad)


def test_refined_reciprocal_is_correctly_rounded(gpu):
    """The MI kernels' exact quotients take one correction step from the
    reciprocal refined by one Newton step, which is IEEE division only if that
    reciprocal is the correctly rounded 1/d: checked on this device for every
    binary32 mantissa of [1, 2) (the steps scale exactly with the exponent)."""
    import ctypes
    n = ctypes.c_uint32(123)
    L.check(L.lib().cgs_check_rcp(ctypes.byref(n), None))
    assert n.value == 0, n.value


def test_mi32_full_and_histograms(gpu, oracle):
    m, f, _ = synth.make_pair(31, 80, sigma=3.0, perturb=4.0, dtype=np.uint8, gamma=True, noise=4.0)
    s, e, c = synth.grid_ranges(3, 4.0)
    fn = A.calcMIstream.with_(bins=32)
    got, (v, i), grid = gpu_search(fn, m, f, s, e, c)
    want = oracle_search(oracle, fn, m, f, s, e, c)
    ok, bad = close(got, want, MI_RTOL)
    assert ok, (bad, got[bad], want[bad])
    vo, io = oracle.argmin(want)
    assert abs(want[i] - vo) <= 2 * MI_RTOL * abs(vo)
    gm, gf = A.CudaImage.from_numpy(m), A.CudaImage.from_numpy(f)
    rng = np.random.default_rng(3)
    for k in list(rng.choice(grid.numElements(), 12, replace=False)) + [i]:
        p = grid.getGridPoint(k)
        for bins in (32, 64):
            fb = fn.with_(bins=bins)
            cost, hist = A.eval_point(fb, gm, gf, p)
            oc, _, oh = oracle.cost_point(m, f, fb.cost, fb.sampler, p, bins=bins, want_hist=True)
            assert np.array_equal(hist, oh), (k, bins)
            assert close([cost], [oc], MI_RTOL)[0]


@pytest.mark.parametrize("fw,fh,fdtype,bins", [(300, 149, np.uint8, 32), (200, 63, np.float32, 16),
                                              (256, 6, np.uint8, 8)])
def test_mi_chunk_shapes_and_bin_counts(gpu, oracle, fw, fh, fdtype, bins):
    """The MI sweep on a full 256-column chunk plus a 44-column one (several
    rows per step), on one 200-column chunk, on heights that are not multiples of
    the row step, on uint8 and float fixed images, at 8, 16 and 32 bins (every
    histogram at the 32-bin stride), with 4 and 8 grid points per workgroup:
    costs against the oracle, joint histograms bit-exact."""
    m, f, _ = synth.make_pair(fw * 1000 + fh, fw, fh, sigma=3.0, perturb=4.0, dtype=np.uint8)
    f = f.astype(fdtype)
    s, e, c = synth.grid_ranges(2, 3.0)
    fn = A.calcMIstream.with_(bins=bins)
    grid = A.CudaGrid(8)
    grid.setStartPoint(s)
    grid.setEndPoint(e)
    grid.setNumSamples(c)
    gm, gf = A.CudaImage.from_numpy(m), A.CudaImage.from_numpy(f)
    want = oracle_search(oracle, fn, m, f, s, e, c)
    for gpt in (4, 8):
        got, _ = _search_with(fn, gm, gf, grid, gpt, None)
        ok, bad = close(got, want, MI_RTOL)
        assert ok, (gpt, bad, got[bad], want[bad])
    for k in (0, 77, 255):
        p = grid.getGridPoint(k)
        cost, hist = A.eval_point(fn, gm, gf, p)
        oc, _, oh = oracle.cost_point(m, f, fn.cost, fn.sampler, p, bins=bins, want_hist=True)
        assert np.array_equal(hist, oh), k


CASES = [
    # bilinear SAD runs the fast kernel (v_rcp sample points): the bilinear FP32 bar, 1e-5
    ("sad_noncentred_bilinear_u8", A.averageAbsoluteDifference, np.uint8, 1, NCC_RTOL),
    ("sad_rawH_bilinear_f32", A.averageAbsoluteDifferenceH, np.float32, 1, NCC_RTOL),
    ("sqd_centred_bilinear_u8", A.calcSQD, np.uint8, 1, NCC_RTOL),
    ("ncc_centred_nearest_u8", A.calcNCC.with_(sampler=L.SAMPLER_NEAREST), np.uint8, 1, NCC_RTOL),
    ("mi_plain_64_nearest", A.calcMI.with_(sampler=L.SAMPLER_NEAREST), np.uint8, 1, MI_RTOL),
    ("sad_centred_nearest_rgb", A.centredAbsoluteDifference.with_(sampler=L.SAMPLER_NEAREST), np.uint8, 3, 0.0),
    ("ncc_centred_bilinear_rgb_f32", A.calcNCC, np.float32, 3, NCC_RTOL),
    ("mi_centred_bilinear_rgb", A.calcMIstream.with_(bins=16), np.uint8, 3, MI_RTOL),
]


@pytest.mark.parametrize("name,fn,dtype,channels,rtol", CASES, ids=[c[0] for c in CASES])
def test_functor_modes(gpu, oracle, name, fn, dtype, channels, rtol):
    rng = np.random.default_rng(hash(name) % 2 ** 32)
    if channels == 1:
        m, f, _ = synth.make_pair(int(rng.integers(1 << 30)), 40, 36, sigma=2.0, perturb=3.0, dtype=dtype)
    else:
        planes = [synth.make_pair(int(rng.integers(1 << 30)), 40, 36, sigma=2.0, perturb=3.0, dtype=dtype)
                  for _ in range(channels)]
        m = np.stack([p[0] for p in planes], -1)
        f = np.stack([p[1] for p in planes], -1)
    if fn.param_mode == L.PARAM_RAW_H:
        s = [0.97, -0.03, -2.0, -0.03, 0.97, -2.0, -4e-4, -4e-4]
        e = [1.03, 0.03, 2.0, 0.03, 1.03, 2.0, 4e-4, 4e-4]
    elif fn.param_mode == L.PARAM_NONCENTRED:
        s = [-0.05, 0.97, 0.97, -0.03, -2.0, -2.0, -2e-4, -2e-4]
        e = [0.05, 1.03, 1.03, 0.03, 2.0, 2.0, 2e-4, 2e-4]
    else:
        s, e, _ = synth.grid_ranges(3, 3.0)
    c = [3] * 8
    got, (v, i), _ = gpu_search(fn, m, f, s, e, c)
    want = oracle_search(oracle, fn, m, f, s, e, c)
    if rtol == 0.0:
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
        assert (v, i) == oracle.argmin(want)
    else:
        ok, bad = close(got, want, rtol, atol=1e-30)
        assert ok, (bad, got[bad], want[bad])
        vo, _ = oracle.argmin(want)
        assert abs(want[i] - vo) <= 2 * rtol * abs(vo)


def test_stream_functor_finish_rules(gpu, oracle):
    """calcNCCstream / calcSQDstream / averageAbsoluteDifference_stream on grids
    whose shifts reach the special cases: no overlap (SQD stream 0/0 = NaN,
    SAD stream 0) and an overlap on the zero half of the fixed image (NCC stream
    0 where calcNCC gives NaN).  Single search and (bilinear cases) the batch
    both match the oracle."""
    m, f, _ = synth.make_pair(44, 40, 36, sigma=2.0, perturb=3.0, dtype=np.uint8)
    f = f.copy()
    f[:, :24] = 0
    s, e, c = synth.grid_ranges(3, 3.0)
    s[4], e[4], c[4] = -60.0, 60.0, 7
    cases = [(A.calcNCCstream, s, e, c), (A.calcSQDstream, s, e, c),
             (A.calcNCCstream.with_(sampler=L.SAMPLER_NEAREST), s, e, c),
             (A.averageAbsoluteDifference_stream, [-50, -50], [50, 50], [21, 21]),
             (A.averageAbsoluteDifference_stream.with_(sampler=L.SAMPLER_NEAREST), [-50, -50], [50, 50], [21, 21])]
    pairs, wants = [], []
    for fn, s_, e_, c_ in cases:
        got, (v, i), grid = gpu_search(fn, m, f, s_, e_, c_)
        want = oracle_search(oracle, fn, m, f, s_, e_, c_)
        exact = fn.cost == L.COST_SAD_STREAM and fn.sampler == L.SAMPLER_NEAREST
        if exact:
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
        else:
            for k in np.where(~close_mask(got, want, NCC_RTOL))[0]:
                assert fn.cost in (L.COST_NCC_STREAM, L.COST_SQD_STREAM), (fn.name, k, got[k], want[k])
                # a point whose few overlapping moving samples are all below one grey
                # level: the fast kernel's absolute sample error (~1e-5 grey) dominates
                # the sum of squares in the denominator; 1e-3 relative applies there
                _, aux = oracle.cost_point(m, f, fn.cost, fn.sampler, grid.getGridPoint(k), mode=fn.param_mode)
                assert aux[2] * 255.0 ** 2 < aux[4] and abs(got[k] - want[k]) <= 1e-3 * abs(want[k]), \
                    (fn.name, k, got[k], want[k], aux[:5])
        vo, io = oracle.argmin(want)
        assert (v, i) == (vo, io) or abs(want[i] - vo) <= 2 * NCC_RTOL * abs(vo)
        # the special cases are reached
        if fn.cost == L.COST_NCC_STREAM:
            plain_ncc = oracle_search(oracle, A.calcNCC.with_(sampler=fn.sampler), m, f, s_, e_, c_)
            assert np.any(np.isnan(plain_ncc) & (want == 0.0))
        elif fn.cost == L.COST_SQD_STREAM:
            assert np.isnan(want).any() and np.isfinite(want).any()
        else:
            assert (want == 0.0).sum() > 1
        if fn.sampler == L.SAMPLER_BILINEAR:   # the batch runs bilinear one-channel searches
            pairs.append((fn, grid, A.CudaImage.from_numpy(m), A.CudaImage.from_numpy(f),
                          A.CudaTensor(grid.getAxisSampleCounts())))
            wants.append(got)
    b = A.BatchSearcher()
    b.plan(pairs)
    b.run(use_graph=False)
    for (fn, _, _, _, res), want in zip(pairs, wants):
        assert np.array_equal(res.data.cpu().numpy().view(np.uint32), want.view(np.uint32)), fn.name


def test_exact_kernel_nonfinite_pixels(gpu, oracle):
    """Three-channel float images (the exact kernel) with a NaN in the moving
    image's last row and an infinity in its last column, under integer
    translations: the bilinear taps at floor and ceil keep them out of the
    samples next to them, as the reference's valueAt does, so the shifts that
    leave that row and column out of the overlap have finite costs."""
    planes = [synth.make_pair(60 + k, 32, 28, sigma=2.0, perturb=3.0, dtype=np.float32) for k in range(3)]
    m = np.stack([p[0] for p in planes], -1)
    f = np.stack([p[1] for p in planes], -1)
    m[27, 12, 1] = np.nan
    m[9, 31, 0] = np.inf
    fn = A.calcNCC.with_(param_mode=L.PARAM_TRANSLATION)
    got, _, _ = gpu_search(fn, m, f, [-2, -2], [2, 2], [5, 5])
    want = oracle_search(oracle, fn, m, f, [-2, -2], [2, 2], [5, 5])
    assert np.array_equal(np.isnan(got), np.isnan(want)) and np.isnan(want).any() and (~np.isnan(want)).any()
    ok = close_mask(got, want, NCC_RTOL)
    assert ok.all(), np.where(~ok)[0][:8]


def test_driver_images_translation(gpu, oracle):
    d = json.load(open(os.path.join(HERE, "golden", "driver_images.json")))
    Aimg = np.array(d["imageA"], np.uint8).reshape(6, 6) * 255
    Cimg = np.array(d["imageC"], np.uint8).reshape(6, 6) * 255
    for fn in (A.sumOfAbsoluteDifferences.with_(sampler=L.SAMPLER_NEAREST), A.sumOfAbsoluteDifferences):
        got, (v, i), _ = gpu_search(fn, Aimg, Cimg, [-5, -5], [5, 5], [21, 21])
        want = oracle_search(oracle, fn, Aimg, Cimg, [-5, -5], [5, 5], [21, 21])
        assert np.allclose(got, want, rtol=1e-6, atol=1e-3)
        assert oracle.argmin(want)[1] == i or abs(want[i] - oracle.argmin(want)[0]) < 1e-3


def test_find_extrema_matches_oracle(gpu, oracle):
    import torch
    rng = np.random.default_rng(2)
    for n in [1, 7, 512, 1025, 100000]:
        v = rng.integers(0, 5, n).astype(np.float32)
        v[rng.integers(0, n, max(1, n // 20))] = np.nan
        t = A.CudaTensor([n], data=torch.from_numpy(v).cuda())
        assert t.find_extrema(L.TIE_LOWEST) == oracle.argmin(v)
        assert t.find_extrema(L.TIE_REFERENCE) == oracle.argmin(v, oracle.TIE_REFERENCE)
    t = A.CudaTensor([3], data=torch.tensor([math.nan, math.inf, math.nan], device="cuda"))
    assert t.find_extrema() == (math.inf, -1)


def test_invalid_arguments_raise(gpu):
    m = A.CudaImage.from_numpy(np.zeros((8, 8), np.uint8))
    grid = A.CudaGrid(2)
    grid.setNumSamples([2, 2])
    res = A.CudaTensor(grid.getAxisSampleCounts())
    with pytest.raises(L.CgsError):
        A.CudaGridSearcher(grid, res).search_by_value(A.calcNCC, m, m)   # 2-D grid, 8-D functor
    with pytest.raises(L.CgsError):
        A.CudaGridSearcher(grid, res).search_by_value(A.calcMIstream.with_(bins=100), m, m)


COV_RTOL = 1e-5   # binary64 sums in another order; log from ocml vs libm, both rounded to binary32


@pytest.mark.parametrize("name", ["ncc", "sad"])
def test_covariance_by_value(gpu, oracle, name):
    """covariance_by_value over a search's cost tensor against the oracle's
    restatement of calculateCovariance_by_value; SAD costs are positive, so
    the reference's scan keeps index 0 / value 0 and the matrix is zero."""
    m, f, _ = synth.make_pair(61, 48, sigma=2.0, perturb=2.0, dtype=np.float32)
    s, e, c = synth.grid_ranges(4, 3.0)
    fn = A.calcNCC if name == "ncc" else A.centredAbsoluteDifference
    got, (v, i), grid = gpu_search(fn, m, f, s, e, c)
    res = A.CudaTensor(grid.getAxisSampleCounts())
    searcher = A.CudaGridSearcher(grid, res)
    searcher.search_by_value(fn, A.CudaImage.from_numpy(m), A.CudaImage.from_numpy(f))
    base = np.eye(8, dtype=np.float32)
    cov = searcher.covariance_by_value(base.copy())
    c2, r = oracle.grid_update(s, e, c)
    want, mi, mv = oracle.covariance(s, r, c2, res.data.cpu().numpy())
    scale = np.abs(want).max()
    if name == "ncc":
        assert mi == i and scale > 0
    else:
        assert mi == 0 and mv == 0.0 and scale == 0
    assert np.allclose(cov - base, want, rtol=COV_RTOL, atol=COV_RTOL * scale)
