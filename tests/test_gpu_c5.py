"""The learned proposal (C5) on 128 held-out 256x256 pairs with the committed
network (cugridsearch_amd/weights/proposal_c5.npz) and its calibrated
threshold and descent starts: argmin agreement with the plain full 6^8 NCC
grid (the arbiter), fallbacks equal to the full grid's answer bit for bit,
local answers carrying the full grid's cost of that point, and what the
network gains over the same refinement from as many fixed starts."""
import os

import numpy as np
import pytest

from cugridsearch_amd import api as A
from cugridsearch_amd import cnn
from cugridsearch_amd import proposal as P

pytestmark = pytest.mark.gpu

WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "cugridsearch_amd", "weights",
                       "proposal_c5.npz")
N = 128


def test_c5_learned_proposal_against_full_grid(gpu):
    net = cnn.ProposalNet.load(WEIGHTS, max_batch=256)
    threshold, modes = P.saved_settings(WEIGHTS)
    assert np.isfinite(threshold) and len(modes) >= 1
    g = P.grid()
    images = []
    for k in range(N):
        m, f = P.pair(k, heldout=True)
        images.append((A.CudaImage.from_numpy(m), A.CudaImage.from_numpy(f)))
    full = [P.full_grid_argmin(g, m, f) for m, f in images]
    ref = np.array([i for _, i in full], np.int32)
    res, _, _ = P.run_learned(net, images, g, threshold, modes=modes)
    agree = float((res["index"] == ref).mean())
    assert agree >= 0.98, agree          # 99.8% measured on 1024 held-out pairs
    fb = res["outcome"] != 0
    for k in np.where(fb)[0]:            # a fallback searches the full grid: the arbiter's answer exactly
        assert res["index"][k] == ref[k] and np.float32(res["value"][k]) == np.float32(full[k][0]), k
    for k in np.where(~fb & (res["index"] == ref))[0]:   # a local answer carries the full grid's cost
        assert np.float32(res["value"][k]).view(np.uint32) == np.float32(full[k][0]).view(np.uint32), k
    assert res["points"].mean() < 0.1 * g.numElements()
    abl, _ = P.run_without_network(images, g, threshold, starts=len(modes))
    agree_abl = float((abl["index"] == ref).mean())
    # the network's gain: better agreement, or the same with fewer boxes searched
    assert agree > agree_abl or (agree == agree_abl and res["steps"].mean() < abl["steps"].mean()), \
        (agree, agree_abl, res["steps"].mean(), abl["steps"].mean())
    print(f"C5 128 held-out pairs: learned {agree:.4f} ({res['steps'].mean():.2f} boxes/pair, "
          f"{fb.mean():.3f} fallbacks), fixed starts {agree_abl:.4f} ({abl['steps'].mean():.2f} boxes/pair)")
