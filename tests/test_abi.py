"""CPU checks of the C-ABI library: it loads, exports every symbol the header
declares, and its host-only entry points (grid bookkeeping, validation)
agree with the oracle.  No GPU compute is called."""
import ctypes
import os
import re

import numpy as np
import pytest

from cugridsearch_amd import _lib as L
from cugridsearch_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    syms = set()
    for name in os.listdir(os.path.join(ROOT, "include")):
        if name.endswith(".h"):
            text = open(os.path.join(ROOT, "include", name)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            syms |= set(re.findall(r"\b(cgs_[a-z0-9_]+)\s*\(", text))
    return syms


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    syms = header_symbols()
    assert "cgs_search" in syms and "cgs_grid_init" in syms
    for s in sorted(syms):
        assert hasattr(lib, s), s
    assert set(L.SIGNATURES) <= syms


def test_version_names_this_trees_sources():
    """The library reports the hash of the sources it was built from, and it
    is this tree's (build() after every source change)."""
    from cugridsearch_amd import build as B
    v = L.lib().cgs_version().decode()
    assert v.startswith("cgs") and v.endswith(" src " + B.source_hash()), (v, B.source_hash())


def _grid(dim, start, end, count):
    g = L.Grid()
    st = L.lib().cgs_grid_init(ctypes.byref(g), dim, L.floats(start), L.floats(end), L.floats(count))
    return st, g


@pytest.mark.parametrize("levels,translate", [(4, 4.0), (5, 8.0), (6, 8.0)])
def test_grid_init_matches_oracle(oracle, levels, translate):
    s, e, c = synth.grid_ranges(levels, translate)
    st, g = _grid(8, s, e, c)
    assert st == 0
    c2, r = oracle.grid_update(s, e, c)
    assert np.array_equal(np.array(g.resolution[:8], np.float32), r)
    assert g.num_elements == oracle.grid_num_elements(c2) == levels ** 8
    out = (ctypes.c_float * 8)()
    rng = np.random.default_rng(0)
    for idx in rng.integers(0, levels ** 8, 200):
        assert L.lib().cgs_grid_point(ctypes.byref(g), int(idx), out) == 0
        assert np.array_equal(np.array(out[:], np.float32), oracle.grid_point(s, r, c2, idx))


def test_grid_init_quirks_match_oracle(oracle):
    cases = [([1.0, 0.0], [5.0, 0.0], [1.0, 7.0]), ([0.0, 0.0], [1.0, 2.0], [2.5, 3.0]),
             ([0.0, 0.0], [3.0, 1.0], [3.0, 2.5])]
    for s, e, c in cases:
        st, g = _grid(2, s, e, c)
        assert st == 0
        c2, r = oracle.grid_update(s, e, c)
        assert np.array_equal(np.array(g.count[:2], np.float32), c2)
        assert np.array_equal(np.array(g.resolution[:2], np.float32), r)
        assert g.num_elements == oracle.grid_num_elements(c2)


def test_grid_init_errors():
    assert _grid(2, [0, 0], [1, 1], [65536, 65536])[0] == 2   # CGS_ERR_GRID_TOO_LARGE
    assert _grid(9, [0] * 9, [1] * 9, [2] * 9)[0] == 1         # bad dimension
    assert _grid(1, [0], [1], [0.5])[0] == 1                   # count < 1
    assert _grid(1, [0], [float("nan")], [2])[0] == 1
    g = L.Grid()
    assert L.lib().cgs_grid_init(ctypes.byref(g), 2, None, None, None) == 1
    assert L.lib().cgs_last_error().decode() != ""
    st, g = _grid(1, [0], [1], [4])
    out = (ctypes.c_float * 1)()
    assert L.lib().cgs_grid_point(ctypes.byref(g), 4, out) == 1   # index out of range


def test_search_rejects_bad_arguments_before_touching_a_device():
    st, g = _grid(8, *synth.grid_ranges(2, 1.0))
    o = L.SearchOpts()
    L.lib().cgs_default_opts(ctypes.byref(o))
    img = L.Image(None, 8, 8, 1, L.PIX_U8)
    assert L.lib().cgs_search(ctypes.byref(g), ctypes.byref(img), ctypes.byref(img), ctypes.byref(o), None, None,
                              None) == 1


def test_line_traversal_options_are_validated():
    """The line-traversal costs take the bilinear sampler and their own
    parametrisation (checked before any device call)."""
    st, g = _grid(8, *synth.grid_ranges(2, 1.0))
    buf = (ctypes.c_uint8 * 64)()
    img = L.Image(ctypes.cast(buf, ctypes.c_void_p), 8, 8, 1, L.PIX_U8)

    def search(cost, sampler=L.SAMPLER_BILINEAR, mode=L.PARAM_CENTRED):
        o = L.SearchOpts()
        L.lib().cgs_default_opts(ctypes.byref(o))
        o.cost, o.sampler, o.param_mode, o.mi_bins = cost, sampler, mode, 64
        return L.lib().cgs_search(ctypes.byref(g), ctypes.byref(img), ctypes.byref(img), ctypes.byref(o), None,
                                  None, None), L.lib().cgs_last_error().decode()

    for cost in (L.COST_SQD_ALT, L.COST_NCC_ALT, L.COST_NCC_STREAM_ALT, L.COST_MI_STREAM_ALT):
        assert search(cost, sampler=L.SAMPLER_NEAREST)[0] == 3                 # CGS_ERR_UNSUPPORTED
        assert search(cost, mode=L.PARAM_NORM)[0] == 1
        assert search(cost, mode=L.PARAM_NONCENTRED)[0] == 1
    assert search(L.COST_MI_ALT, mode=L.PARAM_CENTRED)[0] == 1
    st, msg = search(L.COST_NCC, mode=L.PARAM_NORM)
    assert st == 1 and "calcMIAlt" in msg
    assert search(14)[0] == 1
