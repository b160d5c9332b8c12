"""Tile autotuner: times candidate tile shapes on the device and persists the
best per (width, height, pixel type, cost, sampler) for single-pair searches,
and per batch signature for the persistent batch kernel.

Tile shapes never change results (every grid point's arithmetic and every
reduction order are independent of how many points share a workgroup), so a
tuned and a default search give bit-identical costs; the tests check that.

Choices live in a JSON file: $CGS_TUNING_FILE, else
~/.cache/cugridsearch_amd/tuning.json.
"""
import ctypes
import json
import os
import time

from . import _lib as L

PLAIN_CANDIDATES = {"fast": (4, 6, 8), "mi": (2, 4, 6, 8)}
BATCH_CANDIDATES = [dict(fast_gpt=4, mi_gpt=4, wg_per_cu=4, items_per_claim=8),
                    dict(fast_gpt=6, mi_gpt=4, wg_per_cu=4, items_per_claim=8),
                    dict(fast_gpt=6, mi_gpt=8, wg_per_cu=4, items_per_claim=8),
                    dict(fast_gpt=4, mi_gpt=2, wg_per_cu=4, items_per_claim=16),
                    dict(fast_gpt=6, mi_gpt=8, wg_per_cu=8, items_per_claim=4)]


def cache_path():
    return os.environ.get("CGS_TUNING_FILE",
                          os.path.join(os.path.expanduser("~"), ".cache", "cugridsearch_amd", "tuning.json"))


_cache = {"path": None, "data": None}


def load():
    p = cache_path()
    if _cache["path"] != p or _cache["data"] is None:
        try:
            with open(p) as f:
                _cache["data"] = json.load(f)
        except (OSError, ValueError):
            _cache["data"] = {}
        _cache["path"] = p
    return dict(_cache["data"])


def save(data):
    p = cache_path()
    os.makedirs(os.path.dirname(p), exist_ok=True)
    tmp = p + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, p)
    _cache["path"], _cache["data"] = p, dict(data)


def plain_key(functor, moved, fixed):
    pix = "u8" if moved._img.pixtype == L.PIX_U8 else "f32"
    return f"plain:{fixed.width()}x{fixed.height()}:{pix}:cost{functor.cost}:sampler{functor.sampler}"


def family(functor):
    if functor.sampler != L.SAMPLER_BILINEAR or functor.cost >= L.COST_SQD_ALT:   # line traversal: no tiles
        return None
    return "mi" if functor.cost in (L.COST_MI, L.COST_MI_PLAIN) else "fast"


def tuned_gpt(functor, moved, fixed):
    """The persisted tile choice for this search, or 0 (library default)."""
    return int(load().get(plain_key(functor, moved, fixed), 0))


def tune_plain(functor, moved, fixed, grid, max_points=65536, repeats=2, stream=None):
    """Times each candidate tile size on the first max_points grid points and
    persists the fastest.  Returns {gpt: ms} and the choice."""
    from .api import CudaTensor, _stream_ptr, make_opts
    fam = family(functor)
    if fam is None:
        return {}, 0
    import numpy as np
    sub = type(grid)(grid.getDimension())
    counts = grid.getAxisSampleCounts().copy()
    # shrink the slowest axes until the sub-grid has <= max_points points
    start, end = grid.start_point(), grid.end_point()
    for a in range(len(counts) - 1, -1, -1):
        while np.prod(counts) > max_points and counts[a] > 1:
            counts[a] -= 1
    sub.setStartPoint(start)
    sub.setEndPoint(end)
    sub.setNumSamples(counts)
    res = CudaTensor(sub.getAxisSampleCounts())
    times = {}
    for gpt in PLAIN_CANDIDATES[fam]:
        if fam == "mi" and functor.bins > 32 and gpt > 2:
            continue
        o = make_opts(functor)
        o.tile_gpt = gpt
        o.flags = L.FLAG_TIME_KERNEL
        out = L.Argmin()
        best = float("inf")
        for _ in range(repeats):
            L.check(L.lib().cgs_search(ctypes.byref(sub._g), ctypes.byref(moved._img), ctypes.byref(fixed._img),
                                       ctypes.byref(o), ctypes.c_void_p(res.data.data_ptr()), ctypes.byref(out),
                                       _stream_ptr(stream)))
            best = min(best, L.lib().cgs_last_kernel_ms())
        times[gpt] = best
    choice = min(times, key=times.get)
    data = load()
    data[plain_key(functor, moved, fixed)] = choice
    save(data)
    return times, choice


def batch_key(pairs):
    n = len(pairs)
    pe = sum(g.numElements() * f.width() * f.height() for _, g, _, f, _ in pairs)
    kinds = sorted({family(fn) for fn, *_ in pairs})
    return f"batch:n{n}:pe{int(pe):d}:{'-'.join(kinds)}"


def tuned_batch(pairs):
    return load().get(batch_key(pairs))


def tune_batch(pairs, sample_pairs=24, stream=None):
    """Times each candidate batch tile configuration on a sub-batch of the
    pairs (every kind represented) and persists the fastest for this batch
    signature.  Returns ({index: ms}, choice)."""
    from .api import BatchSearcher
    import torch
    sub = pairs[:sample_pairs]
    times = {}
    s = stream or torch.cuda.current_stream()
    for k, cand in enumerate(BATCH_CANDIDATES):
        b = BatchSearcher()
        b.plan([(fn, g, m, f, None) for fn, g, m, f, _ in sub], tiles=cand)
        b.run(use_graph=False, stream=s)
        b.results(stream=s)
        t0 = time.perf_counter()
        b.run(use_graph=False, stream=s)
        b.results(stream=s)
        times[k] = (time.perf_counter() - t0) * 1000.0
    best = min(times, key=times.get)
    choice = BATCH_CANDIDATES[best]
    data = load()
    data[batch_key(pairs)] = choice
    save(data)
    return times, choice
