"""The ragged batch workload (256 seeded pairs, sides U{64..512}, uint8 and
float32 alternating, 4^8..6^8 grids, SAD:NCC:MI 1:1:1) on its first 24
pairs: sides up to 511, 6^8 grids, every cost kind.  The batch runs with
autotuned tiles from hipGraph replays; every pair's cost tensor and argmin
equal the plain single-pair search bit for bit, and every argmin's cost is
the oracle's cost of that grid point."""
import numpy as np
import pytest

from cugridsearch_amd import api as A
from cugridsearch_amd import synth, tuner

pytestmark = pytest.mark.gpu

FNS = {"sad": A.centredAbsoluteDifference, "ncc": A.calcNCC, "mi": A.calcMIstream.with_(bins=32)}
RTOL = {"sad": 1e-5, "ncc": 1e-5, "mi": 1e-6}


@pytest.fixture(scope="module")
def c4_pairs():
    specs = synth.batch_workload()[:24]
    pairs = []
    for sp in specs:
        g = A.CudaGrid(8)
        g.setStartPoint(sp["start"])
        g.setEndPoint(sp["end"])
        g.setNumSamples(sp["count"])
        pairs.append((FNS[sp["cost"]], g, A.CudaImage.from_numpy(sp["moving"]), A.CudaImage.from_numpy(sp["fixed"]),
                      A.CudaTensor(g.getAxisSampleCounts())))
    return specs, pairs


def test_c4_batch_equals_plain_and_oracle(gpu, oracle, c4_pairs, tmp_path, monkeypatch):
    import torch
    specs, pairs = c4_pairs
    assert max(max(s["moving"].shape) for s in specs) >= 500 and max(s["levels"] for s in specs) == 6
    assert {s["cost"] for s in specs} == {"sad", "ncc", "mi"}
    assert {s["moving"].dtype for s in specs} == {np.dtype(np.uint8), np.dtype(np.float32)}
    monkeypatch.setenv("CGS_TUNING_FILE", str(tmp_path / "tuning.json"))
    stream = torch.cuda.Stream()
    _, tiles = tuner.tune_batch(pairs, sample_pairs=6, stream=stream)
    assert tuner.tuned_batch(pairs) == tiles
    b = A.BatchSearcher()
    b.plan([(fn, g, m, f, res) for fn, g, m, f, res in pairs], tiles=tiles)
    for _ in range(2):   # capture, then a replay
        for *_, res in pairs:
            res.data.fill_(float("nan"))
        torch.cuda.synchronize()
        b.run(use_graph=True, stream=stream)
        got = b.results(stream=stream)
    torch.cuda.synchronize()
    for k, ((fn, g, m, f, res), sp) in enumerate(zip(pairs, specs)):
        plain = A.CudaTensor(g.getAxisSampleCounts())
        want = A.CudaGridSearcher(g, plain).search_by_value(fn, m, f)
        assert got[k] == want, (k, got[k], want)
        assert np.array_equal(res.data.cpu().numpy().view(np.uint32), plain.data.cpu().numpy().view(np.uint32)), k
        v, i = got[k]
        ov, _ = oracle.cost_point(sp["moving"], sp["fixed"], fn.cost, fn.sampler, g.getGridPoint(i), bins=fn.bins)
        assert abs(v - ov) <= RTOL[sp["cost"]] * abs(ov), (k, sp["cost"], v, ov)
