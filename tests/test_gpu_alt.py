"""GPU parity of the line-traversal ("Alt") costs with the oracle, through
the C-ABI.

Bars: calcSQDAlt / calcNCCAlt / calcNCCstreamAlt costs and argmin bit-exact
(the kernel's lane partition and butterfly are the oracle's declared sums);
calcMIstreamAlt / calcMIAlt joint histograms bit-exact and costs within
1e-6 relative (binary64 entropy with device vs libm log2)."""
import numpy as np
import pytest

from cugridsearch_amd import synth
from cugridsearch_amd import api as A
from cugridsearch_amd import _lib as L

pytestmark = pytest.mark.gpu

MI_RTOL = 1e-6
EXACT = ("calcSQDAlt", "calcNCCAlt", "calcNCCstreamAlt")


def run(fn, m, f, s, e, c):
    grid = A.CudaGrid(8)
    grid.setStartPoint(s)
    grid.setEndPoint(e)
    grid.setNumSamples(c)
    res = A.CudaTensor(grid.getAxisSampleCounts())
    v, i = A.CudaGridSearcher(grid, res).search_by_value(fn, A.CudaImage.from_numpy(m), A.CudaImage.from_numpy(f))
    return res.data.cpu().numpy(), (v, i), grid


def want_of(oracle, fn, m, f, s, e, c):
    c2, r = oracle.grid_update(s, e, c)
    return oracle.search(m, f, fn.cost, fn.sampler, s, r, c2, mode=fn.param_mode, bins=fn.bins)


def check(name, got, want, v, i, oracle):
    if name in EXACT:
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), \
            (name, np.where(got.view(np.uint32) != want.view(np.uint32))[0][:8])
        assert (v, i) == oracle.argmin(want)
    else:
        both_nan = np.isnan(got) & np.isnan(want)
        ok = both_nan | (got == want) | (np.abs(got - want) <= MI_RTOL * np.abs(want))
        assert ok.all(), (name, np.where(~ok)[0][:8], got[~ok][:4], want[~ok][:4])
        vo, _ = oracle.argmin(want)
        assert abs(want[i] - vo) <= 2 * MI_RTOL * abs(vo)


def grid_for(fn, levels=3, translate=3.0):
    s, e, c = synth.grid_ranges(levels, translate)
    if fn.cost == L.COST_MI_ALT:   # calcMIAlt's identity has no qualifying row: step off it
        s[4], e[4] = -2.7, 3.3
    return s, e, c


FUNCTORS = [A.calcSQDAlt, A.calcNCCAlt, A.calcNCCstreamAlt, A.calcMIstreamAlt, A.calcMIAlt]


@pytest.mark.parametrize("fn", FUNCTORS, ids=[f.name for f in FUNCTORS])
@pytest.mark.parametrize("dtype,channels", [(np.uint8, 1), (np.float32, 1), (np.uint8, 3)])
def test_alt_small_grid_matches_oracle(gpu, oracle, fn, dtype, channels):
    rng = np.random.default_rng(100 + fn.cost * 7 + channels)
    planes = [synth.make_pair(int(rng.integers(1 << 30)), 40, 36, sigma=2.0, perturb=3.0, dtype=dtype)
              for _ in range(channels)]
    m = planes[0][0] if channels == 1 else np.stack([p[0] for p in planes], -1)
    f = planes[0][1] if channels == 1 else np.stack([p[1] for p in planes], -1)
    s, e, c = grid_for(fn)
    got, (v, i), _ = run(fn, m, f, s, e, c)
    want = want_of(oracle, fn, m, f, s, e, c)
    check(fn.name, got, want, v, i, oracle)


@pytest.mark.parametrize("fn", FUNCTORS, ids=[f.name for f in FUNCTORS])
def test_alt_sentinels_and_sizes(gpu, oracle, fn):
    """Unequal image sizes (the reference's f_height / m_width quirks bite),
    shifts large enough for the overlap test to fail, and a 3-level grid.
    calcMIAlt gets a moving image narrower than the fixed one: the other way
    round its integer size ratio is 0 and every cost is -0."""
    mw, mh = (40, 22) if fn.cost == L.COST_MI_ALT else (52, 30)
    m, _, _ = synth.make_pair(7, mw, mh, sigma=2.0, perturb=3.0, dtype=np.uint8)
    _, f, _ = synth.make_pair(8, 44, 48, sigma=2.0, perturb=3.0, dtype=np.uint8)
    s, e, c = synth.grid_ranges(3, 3.0)
    s[4], e[4], c[4] = -40.0, 40.0, 5
    got, (v, i), _ = run(fn, m, f, s, e, c)
    want = want_of(oracle, fn, m, f, s, e, c)
    check(fn.name, got, want, v, i, oracle)
    if fn.cost != L.COST_MI_ALT:
        sentinel = 0.0 if fn.cost == L.COST_NCC_STREAM_ALT else 123123.0
        assert (want == sentinel).sum() > 0 and (want != sentinel).sum() > 0
    else:
        assert (want < 0).all()


@pytest.mark.parametrize("size", [(48, 40), (64, 64), (64, 40)], ids=["48x40", "64x64", "64x40"])
@pytest.mark.parametrize("fn", [A.calcMIstreamAlt, A.calcMIAlt], ids=["calcMIstreamAlt", "calcMIAlt"])
def test_alt_mi_histograms_bit_exact(gpu, oracle, fn, size):
    """Joint histograms equal the oracle's count for count.  64x64 takes
    calcMIAlt's binary32 rescale (power-of-two moving sizes); 48x40 and
    64x40 its binary64 one."""
    m, f, _ = synth.make_pair(9, size[0], size[1], sigma=2.0, perturb=3.0, dtype=np.uint8, gamma=True, noise=4.0)
    gm, gf = A.CudaImage.from_numpy(m), A.CudaImage.from_numpy(f)
    rng = np.random.default_rng(4)
    for _ in range(10):
        p = [rng.uniform(-0.08, 0.08), rng.uniform(0.95, 1.05), rng.uniform(0.95, 1.05), rng.uniform(-.04, .04),
             rng.uniform(-4, 4), rng.uniform(-4, 4), rng.uniform(-1e-3, 1e-3), rng.uniform(-1e-3, 1e-3)]
        for bins in (32, 64):
            fb = fn.with_(bins=bins)
            cost, hist = A.eval_point(fb, gm, gf, p)
            oc, _, oh = oracle.cost_point(m, f, fb.cost, fb.sampler, np.float32(p), bins=bins, want_hist=True)
            assert np.array_equal(hist, oh), (p, bins)
            assert cost == oc or abs(cost - oc) <= MI_RTOL * abs(oc), (p, bins, cost, oc)


def test_alt_ncc_256_sample(gpu, oracle):
    """The headline image size: 256 x 256 float images, a 3^8 grid, bit-exact
    costs on a sample of points and at the argmin."""
    m, f, s, e, c, spec = synth.workload("bilinear_ncc_256")
    s, e, c = synth.grid_ranges(3, 4.0)
    got, (v, i), grid = run(A.calcNCCAlt, m, f, s, e, c)
    rng = np.random.default_rng(6)
    for k in list(rng.choice(grid.numElements(), 24, replace=False)) + [i]:
        w, _ = oracle.cost_point(m, f, A.calcNCCAlt.cost, L.SAMPLER_BILINEAR, grid.getGridPoint(k))
        assert np.float32(w).view(np.uint32) == got[k].view(np.uint32), (k, w, got[k])
    assert v == got.min() and got[i] == v


def test_alt_batch_refused(gpu):
    m, f, _ = synth.make_pair(3, 32, dtype=np.float32)
    s, e, c = synth.grid_ranges(2, 1.0)
    grid = A.CudaGrid(8)
    grid.setStartPoint(s)
    grid.setEndPoint(e)
    grid.setNumSamples(c)
    b = A.BatchSearcher()
    with pytest.raises(Exception, match="line-traversal"):
        b.plan([(A.calcNCCAlt, grid, A.CudaImage.from_numpy(m), A.CudaImage.from_numpy(f), None)])


@pytest.mark.parametrize("fn", [A.calcNCCAlt, A.calcMIstreamAlt], ids=["calcNCCAlt", "calcMIstreamAlt"])
def test_alt_nonfinite_fixed_image(gpu, oracle, fn):
    """A NaN and an infinity in the fixed image.  valueAt reads its taps at
    floor and ceil, so at the integral row (and integral lambda) a non-finite
    neighbour does not leak in, while an infinite pixel itself gives
    1 * inf + 0 * inf = NaN; the kernels follow that bit for bit (MI bins a
    NaN into bin 0)."""
    m, f, _ = synth.make_pair(12, 40, 36, sigma=2.0, perturb=3.0, dtype=np.float32)
    f = f.copy()
    f[20, 11] = np.nan
    f[7, 30] = np.inf
    s, e, c = grid_for(fn)
    got, _, _ = run(fn, m, f, s, e, c)
    want = want_of(oracle, fn, m, f, s, e, c)
    assert np.array_equal(np.isnan(got), np.isnan(want))
    assert np.isnan(want).any() or fn.cost != L.COST_NCC_ALT   # MI bins a NaN into bin 0
    fin = ~np.isnan(want)
    if fn.cost == L.COST_NCC_ALT:
        assert np.array_equal(got[fin].view(np.uint32), want[fin].view(np.uint32))
    else:
        assert np.all((got[fin] == want[fin]) | (np.abs(got[fin] - want[fin]) <= MI_RTOL * np.abs(want[fin])))


CASES_256 = [(fn, np.float32, 256, 256, 256, 256) for fn in FUNCTORS] + \
    [(fn, np.uint8, 256, 256, 256, 256) for fn in FUNCTORS] + [(A.calcMIAlt, np.uint8, 200, 168, 200, 168)] + \
    [(A.calcMIAlt, np.uint8, 512, 512, 512, 512), (A.calcMIAlt, np.uint8, 300, 150, 600, 300),
     (A.calcMIAlt, np.float32, 512, 512, 512, 512)]


@pytest.mark.parametrize("fn,dtype,mw,mh,fw,fh", CASES_256,
                         ids=[f"{c[0].name}-{np.dtype(c[1]).name}-{c[2]}x{c[3]}" for c in CASES_256])
def test_alt_256_sample(gpu, oracle, fn, dtype, mw, mh, fw, fh):
    """Rows of 256 samples take four 64-step passes (the second and later
    passes start from the previous pass's last lambda), 256 rows make four
    64-row blocks over the four wavefronts (MI: one shared LDS histogram), and
    200 x 168 is a moving size whose calcMIAlt rescale takes the binary64
    form.  calcMIAlt at 512 x 512 (the benchmark's size: rows of 511 steps of
    1/511 in 8 passes, the short segments near lambda = 0 sharing the first)
    and with a 600-wide fixed image twice the moving one's width.  A 3^8 grid; 24 sampled points and the argmin against the oracle:
    SQD/NCC costs bit for bit, MI joint histograms bit for bit and costs within
    1e-6."""
    m, _, _ = synth.make_pair(70 + fn.cost, mw, mh, sigma=4.0, perturb=6.0, dtype=dtype, gamma=dtype == np.uint8,
                              noise=3.0 if dtype == np.uint8 else 0.0)
    _, f, _ = synth.make_pair(70 + fn.cost, fw, fh, sigma=4.0, perturb=6.0, dtype=dtype)
    s, e, c = grid_for(fn, levels=3, translate=6.0)
    got, (v, i), grid = run(fn, m, f, s, e, c)
    gm, gf = A.CudaImage.from_numpy(m), A.CudaImage.from_numpy(f)
    rng = np.random.default_rng(fn.cost + mw)
    nonempty = 0
    for k in list(rng.choice(grid.numElements(), 24, replace=False)) + [i]:
        p = grid.getGridPoint(k)
        if fn.name in EXACT:
            w, _ = oracle.cost_point(m, f, fn.cost, L.SAMPLER_BILINEAR, p, mode=fn.param_mode)
            assert np.float32(w).view(np.uint32) == got[k].view(np.uint32), (k, w, got[k])
        else:
            w, _, oh = oracle.cost_point(m, f, fn.cost, L.SAMPLER_BILINEAR, p, mode=fn.param_mode, bins=fn.bins,
                                         want_hist=True)
            cost, hist = A.eval_point(fn, gm, gf, p)
            assert np.array_equal(hist, oh), k
            nonempty += hist.sum() > 0
            for g in (got[k], cost):
                assert g == w or abs(g - w) <= MI_RTOL * abs(w), (k, g, w)
    assert v == np.nanmin(got) and got[i] == v
    assert fn.name in EXACT or nonempty >= 12, nonempty


@pytest.mark.parametrize("fn", [A.calcSQDAlt, A.calcNCCAlt, A.calcNCCstreamAlt, A.calcMIstreamAlt],
                         ids=["calcSQDAlt", "calcNCCAlt", "calcNCCstreamAlt", "calcMIstreamAlt"])
def test_alt_overlap_test_near_threshold(gpu, oracle, fn):
    """A 36-point x-translation axis across the moved-area threshold of a
    256 x 256 float pair at scale 1.3.  The kernels decide the overlap test
    from a binary64 area and its rounding bound, and form the reference's
    binary32 area where the bound straddles the threshold; costs equal the
    oracle's (its literal binary32 loops) at every point.  The sweep holds
    points of that exact path for every functor and, for the per-step sums of
    calcSQDAlt / calcNCCAlt, points where the binary32 and the exact area
    decide the test differently."""
    m, f, _ = synth.make_pair(77, 256, 256, sigma=4.0, perturb=4.0, dtype=np.float32)
    s = np.float32([0.02, 1.3, 1.3, 0, 186.76, 0.3, 0, 0])
    e = s.copy()
    e[4] = 186.795
    c = np.float32([1, 1, 1, 1, 36, 1, 1, 1])
    got, (v, i), grid = run(fn, m, f, s, e, c)
    want = want_of(oracle, fn, m, f, s, e, c)
    check(fn.name, got, want, v, i, oracle)
    denom = float(np.float32(np.float32(256 * 256) * np.float32(1.3)))
    undecided = flips = 0
    for k in range(grid.numElements()):
        _, aux = oracle.cost_point(m, f, fn.cost, L.SAMPLER_BILINEAR, grid.getGridPoint(k), bins=fn.bins)
        S, nadds = aux[9], aux[7] + 256
        E = nadds * 2.0 ** (np.frexp(S * (1 + nadds * 2.0 ** -23))[1] - 25)
        undecided += (S - E) / denom <= 0.25 < (S + E) / denom
        flips += (float(np.float32(np.float32(aux[6]) / np.float32(denom))) > 0.25) != \
            (float(np.float32(np.float32(S) / np.float32(denom))) > 0.25)
    assert undecided >= 10, undecided
    if fn.cost in (L.COST_SQD_ALT, L.COST_NCC_ALT):
        assert flips >= 10, flips
