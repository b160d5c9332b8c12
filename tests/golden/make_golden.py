"""Writes tests/golden/workloads.json: the oracle's full-grid argmin and a fixed
sample of per-point costs for each named workload (cugridsearch_amd.synth).

Usage: python tests/golden/make_golden.py [workload ...]
The full grids take minutes on a few cores (1e11 pixel evaluations for the
256x256 NCC and 512x512 MI workloads); the result is committed so the GPU
tests and the bench can check the full-size argmin without re-running it.
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from cugridsearch_amd import synth  # noqa: E402
from oracle import oracle as O  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "workloads.json")
COSTS = {"sad": O.COST_SAD, "ncc": O.COST_NCC, "mi": O.COST_MI}
SAMPLERS = {"nearest": O.NEAREST, "bilinear": O.BILINEAR}
SAMPLE = 2048


def run(name):
    m, f, start, end, count, spec = synth.workload(name)
    c2, res = O.grid_update(start, end, count)
    n = O.grid_num_elements(c2)
    t = time.time()
    out = O.search(m, f, COSTS[spec["cost"]], SAMPLERS[spec["sampler"]], start, res, c2, bins=spec.get("bins", 32))
    dt = time.time() - t
    v, i = O.argmin(out)
    vr, ir = O.argmin(out, O.TIE_REFERENCE)
    rng = np.random.default_rng(1234)
    idx = np.sort(rng.choice(n, size=min(SAMPLE, n), replace=False))
    order = np.argsort(out, kind="stable")[:16]
    return {
        "generator_version": synth.GENERATOR_VERSION,
        "num_elements": int(n),
        "fixed_sha256": hashlib.sha256(np.ascontiguousarray(f).tobytes()).hexdigest(),
        "moving_sha256": hashlib.sha256(np.ascontiguousarray(m).tobytes()).hexdigest(),
        "argmin_value": float(v), "argmin_index": int(i),
        "argmin_reference_value": float(vr), "argmin_reference_index": int(ir),
        "best16_index": [int(k) for k in order], "best16_value": [float(out[k]) for k in order],
        "sample_index": [int(k) for k in idx], "sample_value": [float(out[k]) for k in idx],
        "nan_count": int(np.isnan(out).sum()),
        "oracle_seconds": dt, "oracle_threads": O.max_threads(),
    }


def main(names):
    data = json.load(open(OUT)) if os.path.exists(OUT) else {}
    for name in names:
        print("oracle full grid:", name, flush=True)
        data[name] = run(name)
        print(name, data[name]["argmin_index"], data[name]["argmin_value"], data[name]["oracle_seconds"], flush=True)
        json.dump(data, open(OUT, "w"), indent=1)


if __name__ == "__main__":
    main(sys.argv[1:] or list(synth.WORKLOADS))
