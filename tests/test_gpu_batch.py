"""Batched ragged search: every pair's costs and argmin equal the plain
single-pair search bit for bit, for eager launches, hipGraph replays (with
eager work and new image contents between replays) and every tile shape."""
import numpy as np
import pytest
import torch

from cugridsearch_amd import _lib as L
from cugridsearch_amd import api as A
from cugridsearch_amd import synth

pytestmark = pytest.mark.gpu

FUNCTORS = {"sad": A.centredAbsoluteDifference, "ncc": A.calcNCC, "mi": A.calcMIstream.with_(bins=32)}


def make_batch(n=12, seed=11, sides=(40, 96), levels=(3,)):
    specs = synth.batch_workload(n=n, seed=seed, sides=sides, levels=levels)
    pairs = []
    for sp in specs:
        g = A.CudaGrid(8)
        g.setStartPoint(sp["start"])
        g.setEndPoint(sp["end"])
        g.setNumSamples(sp["count"])
        res = A.CudaTensor(g.getAxisSampleCounts())
        pairs.append((FUNCTORS[sp["cost"]], g, A.CudaImage.from_numpy(sp["moving"]),
                      A.CudaImage.from_numpy(sp["fixed"]), res))
    return specs, pairs


def plain(pairs):
    out = []
    for fn, g, m, f, _ in pairs:
        r = A.CudaTensor(g.getAxisSampleCounts())
        am = A.CudaGridSearcher(g, r).search_by_value(fn, m, f)
        out.append((r.data.cpu().numpy(), am))
    return out


def check_equal(pairs, got_argmins, want):
    for (fn, g, m, f, res), am, (wcost, wam) in zip(pairs, got_argmins, want):
        assert np.array_equal(res.data.cpu().numpy().view(np.uint32), wcost.view(np.uint32)), fn.name
        assert am == wam, (fn.name, am, wam)


def test_batch_eager_equals_plain(gpu):
    specs, pairs = make_batch()
    want = plain(pairs)
    b = A.BatchSearcher()
    b.plan(pairs)
    b.run(use_graph=False)
    check_equal(pairs, b.results(), want)
    info = b.info()
    assert info["points"] == sum(g.numElements() for _, g, _, _, _ in pairs)


@pytest.mark.parametrize("tiles", [dict(fast_gpt=4, mi_gpt=2, wg_per_cu=2, items_per_claim=1),
                                   dict(fast_gpt=6, mi_gpt=4, wg_per_cu=8, items_per_claim=64),
                                   dict(fast_gpt=4, mi_gpt=8, wg_per_cu=4, items_per_claim=8)])
def test_batch_tiles_identical(gpu, tiles):
    specs, pairs = make_batch(seed=12)
    want = plain(pairs)
    b = A.BatchSearcher()
    b.plan(pairs, tiles=tiles)
    b.run()
    check_equal(pairs, b.results(), want)


def test_batch_mixed_bin_counts(gpu):
    """MI pairs of 16, 32, 33 and 64 bins in one batch at 2 grid points per MI
    item: up to 32 bins a pair holds two histogram copies, above it one, and
    the launch's LDS holds the largest of them (here a 32-bin pair's two
    copies outgrow a 33-bin pair's one)."""
    specs, pairs = make_batch(n=12, seed=14)
    bins_cycle = [16, 32, 33, 64]
    k = 0
    mixed = []
    for (fn, g, m, f, res), sp in zip(pairs, specs):
        if sp["cost"] == "mi":
            fn = fn.with_(bins=bins_cycle[k % len(bins_cycle)])
            k += 1
        mixed.append((fn, g, m, f, res))
    assert k >= 4
    want = plain(mixed)
    b = A.BatchSearcher()
    b.plan(mixed, tiles=dict(fast_gpt=6, mi_gpt=2, wg_per_cu=4, items_per_claim=4))
    b.run(use_graph=False)
    check_equal(mixed, b.results(), want)


def test_batch_graph_replays(gpu):
    specs, pairs = make_batch(seed=13)
    s = torch.cuda.Stream()
    b = A.BatchSearcher()
    b.plan(pairs)
    want = plain(pairs)
    for rep in range(3):
        b.run(use_graph=True, stream=s)
        got = b.results(stream=s)
        check_equal(pairs, got, want)
        # eager work between replays
        plain(pairs[:2])
        torch.cuda.synchronize()
    # new image contents in the same buffers: the replay restages them
    for k, (fn, g, m, f, res) in enumerate(pairs):
        m.tensor.copy_(torch.flip(m.tensor, dims=[1]))
    torch.cuda.synchronize()
    want2 = plain(pairs)
    b.run(use_graph=True, stream=s)
    check_equal(pairs, b.results(stream=s), want2)


def test_batch_rejects_unsupported(gpu):
    _, pairs = make_batch(n=2)
    fn, g, m, f, r = pairs[0]
    b = A.BatchSearcher()
    with pytest.raises(L.CgsError):
        b.plan([(fn.with_(sampler=L.SAMPLER_NEAREST), g, m, f, r)])


def test_autotuner_tiles_bit_identical(gpu, tmp_path, monkeypatch):
    from cugridsearch_amd import tuner
    monkeypatch.setenv("CGS_TUNING_FILE", str(tmp_path / "tuning.json"))
    specs, pairs = make_batch(n=3, seed=21, sides=(60, 90))
    for fn, g, m, f, res in pairs:
        want = A.CudaTensor(g.getAxisSampleCounts())
        wam = A.CudaGridSearcher(g, want).search_by_value(fn, m, f)   # library default tiles
        times, choice = tuner.tune_plain(fn, m, f, g, max_points=4096)
        assert choice in times and tuner.tuned_gpt(fn, m, f) == choice
        for gpt in times:                                               # every candidate, incl. the choice
            o = A.make_opts(fn)
            o.tile_gpt = gpt
            import ctypes
            out = L.Argmin()
            got = A.CudaTensor(g.getAxisSampleCounts())
            L.check(L.lib().cgs_search(ctypes.byref(g._g), ctypes.byref(m._img), ctypes.byref(f._img),
                                       ctypes.byref(o), ctypes.c_void_p(got.data.data_ptr()), ctypes.byref(out),
                                       None))
            assert np.array_equal(got.data.cpu().numpy().view(np.uint32), want.data.cpu().numpy().view(np.uint32))
            assert (out.value, out.index) == wam
