"""cgs_search_noalloc (the search whose scratch lives in a static device array)
against cgs_search, through the C-ABI: costs bit for bit and the same argmin,
on two calls in a row per kernel family, the second with a smaller input,
plus its refusals (line-traversal costs, a moving image beyond the scratch)."""
import ctypes

import numpy as np
import pytest

from cugridsearch_amd import _lib as L
from cugridsearch_amd import api as A
from cugridsearch_amd import synth

pytestmark = pytest.mark.gpu


def _grid(levels, translate):
    s, e, c = synth.grid_ranges(levels, translate)
    g = A.CudaGrid(8)
    g.setStartPoint(s)
    g.setEndPoint(e)
    g.setNumSamples(c)
    return g


def _run(entry, functor, grid, m, f):
    o = A.make_opts(functor)
    out = L.Argmin()
    res = A.CudaTensor(grid.getAxisSampleCounts())
    res.data.fill_(float("nan"))
    st = getattr(L.lib(), entry)(ctypes.byref(grid._g), ctypes.byref(m._img), ctypes.byref(f._img), ctypes.byref(o),
                                 ctypes.c_void_p(res.data.data_ptr()), ctypes.byref(out), None)
    return st, res.data.cpu().numpy(), (out.value, out.index)


def _pair(seed, side, dtype, **kw):
    m, f, _ = synth.make_pair(seed, side, sigma=3.0, perturb=3.0, dtype=dtype, **kw)
    return A.CudaImage.from_numpy(m), A.CudaImage.from_numpy(f)


CASES = [
    # (name, functor, first (side, dtype, levels), second, smaller)
    ("fast_ncc_f32", A.calcNCC, (256, np.float32, 4), (64, np.float32, 3)),
    ("fast_sad_u8", A.centredAbsoluteDifference, (200, np.uint8, 4), (48, np.uint8, 3)),
    ("mi32_u8", A.calcMIstream.with_(bins=32), (128, np.uint8, 3), (64, np.uint8, 3)),
    ("exact_nn_sad_u8", A.centredAbsoluteDifference.with_(sampler=L.SAMPLER_NEAREST), (96, np.uint8, 4),
     (40, np.uint8, 3)),
]


@pytest.mark.parametrize("name,fn,first,second", CASES, ids=[c[0] for c in CASES])
def test_noalloc_equals_search_twice(gpu, name, fn, first, second):
    for k, (side, dtype, levels) in enumerate((first, second)):
        print(f"noalloc case {name} call {k}: {side}x{side} {np.dtype(dtype).name} {levels}^8", flush=True)
        m, f = _pair(50 + k, side, dtype)
        grid = _grid(levels, 4.0)
        st0, want, wam = _run("cgs_search", fn, grid, m, f)
        st1, got, gam = _run("cgs_search_noalloc", fn, grid, m, f)
        assert st0 == 0 and st1 == 0, (st0, st1, L.lib().cgs_last_error())
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), name
        assert gam == wam


def test_noalloc_refusals_then_works(gpu):
    m, f = _pair(60, 64, np.float32)
    grid = _grid(3, 4.0)
    st, _, _ = _run("cgs_search_noalloc", A.calcNCCAlt, grid, m, f)
    assert st == 3, L.lib().cgs_last_error()          # CGS_ERR_UNSUPPORTED
    assert b"line-traversal" in L.lib().cgs_last_error()
    big_m, big_f = _pair(61, 1024, np.float32)
    st, _, _ = _run("cgs_search_noalloc", A.calcNCC, grid, big_m, big_f)
    assert st == 6, L.lib().cgs_last_error()          # CGS_ERR_TOO_LARGE
    assert b"scratch" in L.lib().cgs_last_error()
    st0, want, wam = _run("cgs_search", A.calcNCC, grid, m, f)
    st1, got, gam = _run("cgs_search_noalloc", A.calcNCC, grid, m, f)
    assert st0 == 0 and st1 == 0
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32)) and gam == wam
