"""Grid sub-boxes and local refinement with full-grid fallback, through the
C-ABI.  Bars: a sub-box search returns the full search's costs bit for bit
at the mapped indices; a refinement that stops inside its box answers a
local minimum of the full cost tensor over its 3^8 neighbourhood, with the
full tensor's cost bit for bit; every fallback answers the full-grid argmin
bit for bit."""
import itertools

import numpy as np
import pytest

from cugridsearch_amd import _lib as L
from cugridsearch_amd import api as A
from cugridsearch_amd import proposal as P
from cugridsearch_amd import synth

pytestmark = pytest.mark.gpu

FUNCTORS = {"sad": A.centredAbsoluteDifference, "ncc": A.calcNCC, "mi": A.calcMIstream.with_(bins=32)}


def full(fn, g, m, f):
    r = A.CudaTensor(g.getAxisSampleCounts())
    v, i = A.CudaGridSearcher(g, r).search_by_value(fn, m, f)
    return r.data.cpu().numpy(), (v, i)


def small_pairs(n=6, seed=21):
    specs = synth.batch_workload(n=n, seed=seed, sides=(40, 72), levels=(4,))
    out = []
    for sp in specs:
        g = A.CudaGrid(8)
        g.setStartPoint(sp["start"])
        g.setEndPoint(sp["end"])
        g.setNumSamples(sp["count"])
        out.append((FUNCTORS[sp["cost"]], g, A.CudaImage.from_numpy(sp["moving"]), A.CudaImage.from_numpy(sp["fixed"])))
    return out


def box_to_full(box, g):
    """Full-grid index of every box point, in box order."""
    cnt = box.getAxisSampleCounts().astype(int)
    off = box.offset()
    full_count = g.getAxisSampleCounts().astype(int)
    inc = np.cumprod(np.r_[1, full_count[:-1]])
    idx = []
    for k in range(box.numElements()):
        ax = P.index_axes(k, cnt)
        idx.append(int(((ax + off) * inc).sum()))
    return np.array(idx)


@pytest.mark.parametrize("kind", ["ncc", "mi", "sad"])
def test_subbox_costs_bit_identical(gpu, kind):
    fn, g, m, f = [p for p in small_pairs(9) if p[0] is FUNCTORS[kind]][0]
    want, _ = full(fn, g, m, f)
    box = g.subbox([1, 0, 2, 1, 0, 1, 2, 0], [2, 3, 2, 3, 1, 2, 2, 4])
    assert box.numElements() == 2 * 3 * 2 * 3 * 1 * 2 * 2 * 4
    got, (bv, bi) = full(fn, box, m, f)
    ref = want[box_to_full(box, g)]
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    assert bv == ref.min() and bi == int(np.argmin(ref))
    for k in (0, 17, box.numElements() - 1):
        assert np.array_equal(box.getGridPoint(k), g.getGridPoint(box_to_full(box, g)[k]))


def neighbourhood_min(cost, count, index):
    ax = P.index_axes(index, count)
    inc = np.cumprod(np.r_[1, np.asarray(count[:-1], int)])
    best = np.inf
    for d in itertools.product((-1, 0, 1), repeat=len(count)):
        q = ax + np.array(d)
        if (q < 0).any() or (q >= count).any():
            continue
        best = min(best, cost[int((q * inc).sum())])
    return best


def test_refine_local_minima_and_fallbacks(gpu):
    pairs = small_pairs(6)
    fulls = [full(fn, g, m, f) for fn, g, m, f in pairs]
    rng = np.random.default_rng(5)
    centres = rng.integers(0, 4, (len(pairs), 8)).astype(np.int32)
    b = A.BatchSearcher()
    res = b.refine([(fn, g, m, f, None) for fn, g, m, f in pairs], centres, radius=1, max_steps=6)
    for (fn, g, m, f), (cost, (fv, fi)), r in zip(pairs, fulls, res):
        count = g.getAxisSampleCounts().astype(int)
        assert r["steps"] >= 1 and r["points"] >= 2 ** 8
        if r["outcome"] == L.REFINE_LOCAL:
            assert r["index"] == r["local_index"]
            assert np.float32(r["value"]).view(np.uint32) == cost[r["index"]].view(np.uint32)
            assert r["value"] <= neighbourhood_min(cost, count, r["index"])
        else:
            assert (r["index"], np.float32(r["value"])) == (fi, np.float32(fv)), fn.name
            assert r["points"] >= g.numElements()


def test_refine_from_the_argmin_stops_there(gpu):
    pairs = small_pairs(4, seed=23)
    fulls = [full(fn, g, m, f) for fn, g, m, f in pairs]
    centres = np.stack([P.index_axes(fi, g.getAxisSampleCounts()) for (_, g, _, _), (_, (_, fi)) in zip(pairs, fulls)])
    res = A.BatchSearcher().refine([(fn, g, m, f, None) for fn, g, m, f in pairs], centres, max_steps=1)
    for r, (_, (fv, fi)) in zip(res, fulls):
        assert r["outcome"] == L.REFINE_LOCAL and r["steps"] == 1
        assert (r["index"], np.float32(r["value"])) == (fi, np.float32(fv))


def test_refine_threshold_forces_fallback(gpu):
    pairs = small_pairs(4, seed=24)
    fulls = [full(fn, g, m, f) for fn, g, m, f in pairs]
    centres = np.full((len(pairs), 8), 1, np.int32)
    res = A.BatchSearcher().refine([(fn, g, m, f, None) for fn, g, m, f in pairs], centres, threshold=-np.inf)
    for r, (_, (fv, fi)) in zip(res, fulls):
        assert r["outcome"] in (L.REFINE_FALLBACK_THRESHOLD, L.REFINE_FALLBACK_BOUNDARY)
        assert (r["index"], np.float32(r["value"])) == (fi, np.float32(fv))
    res = A.BatchSearcher().refine([(fn, g, m, f, None) for fn, g, m, f in pairs], centres, threshold=-np.inf,
                                   fallback=False)
    assert (res["index"] == -1).all()


def test_refine_rejects_bad_options(gpu):
    fn, g, m, f = small_pairs(1)[0]
    b = A.BatchSearcher()
    for kw in (dict(radius=0), dict(max_steps=0), dict(threshold=float("nan"))):
        with pytest.raises(L.CgsError):
            b.refine([(fn, g, m, f, None)], np.zeros((1, 8), np.int32), **kw)


def test_refine_multi_start_keeps_the_best(gpu):
    """Several starts per pair answer the lowest of the local minima the
    same starts reach one at a time (lower index on ties)."""
    pairs = small_pairs(5, seed=25)
    fulls = [full(fn, g, m, f) for fn, g, m, f in pairs]
    rng = np.random.default_rng(9)
    centres = rng.integers(0, 4, (len(pairs), 3, 8)).astype(np.int32)
    arg = [(fn, g, m, f, None) for fn, g, m, f in pairs]
    multi = A.BatchSearcher().refine(arg, centres, max_steps=6, fallback=False)
    singles = [A.BatchSearcher().refine(arg, centres[:, k, :], max_steps=6, fallback=False) for k in range(3)]
    for i, ((fn, g, m, f), (cost, _)) in enumerate(zip(pairs, fulls)):
        found = [(s[i]["value"], s[i]["index"]) for s in singles if s[i]["outcome"] == L.REFINE_LOCAL]
        if not found:
            assert multi[i]["outcome"] == L.REFINE_FALLBACK_BOUNDARY
            continue
        v, idx = min(found)
        assert multi[i]["outcome"] == L.REFINE_LOCAL, fn.name
        assert (np.float32(multi[i]["value"]), multi[i]["index"]) == (np.float32(v), idx)
        assert np.float32(multi[i]["value"]).view(np.uint32) == cost[idx].view(np.uint32)
        assert multi[i]["steps"] <= sum(s[i]["steps"] for s in singles)
