"""Learned proposal for the homography grid search: training data, labels,
proposal -> grid-index conversion and the propose / refine / fall back
pipeline.  No reference counterpart (the reference always searches the full
grid, src/gpu/cudaGridSearch.cuh:736-765); the full grid search stays the
arbiter of every result.

Workload (256x256 FP32 pairs, NCC, centred parametrisation):
  pairs   synth.make_pair(seed, 256, sigma=4, perturb=32, float32,
          extra_shapes=True): blurred noise plus random rectangles and
          half-plane edges; the 4 corners perturbed uniformly by +-32 px;
  grid    8 axes x 6 levels = 1,679,616 points spanning theta +-0.15 rad,
          scale 0.86..1.14, shear +-0.2, translation +-25 px, keystone
          +-1.2e-3 (about the 5th..95th percentiles of the parameters of
          +-32 px corner perturbations);
  labels  the argmin index of the plain full-grid NCC search of each pair.

Command line (labelling runs on the GPU):
  python -m cugridsearch_amd.proposal label --first 0 --count 12000 \
      --seconds 840 --out labels.npz
"""
import argparse
import ctypes
import math
import os
import sys
import time

import numpy as np

from . import synth

SIZE = 256
SIGMA = 4.0
PERTURB = 32.0
LEVELS = 6
GRID_START = np.array([-0.15, 0.86, 0.86, -0.2, -25.0, -25.0, -1.2e-3, -1.2e-3], np.float32)
GRID_END = np.array([0.15, 1.14, 1.14, 0.2, 25.0, 25.0, 1.2e-3, 1.2e-3], np.float32)
GRID_COUNT = np.full(8, LEVELS, np.float32)
TRAIN_SEED = 0xC6505 << 32
HELDOUT_SEED = 0xC6506 << 32
OFFSET_SCALE = 32.0      # corner offsets are regressed in units of 32 px


# --------------------------------------------------------------- geometry

def homography(params, cx, cy):
    """FP64 centred homography of grid parameters (theta, sx, sy, shear, tx,
    ty, kx, ky): Tc K T R Sh Sc Tc^-1 normalised to h33 = 1, with Tc the
    translation by (cx, cy), Sc = diag(sx, sy), Sh the x-shear, R the
    rotation, T the translation and K the keystone row (kx, ky, 1) -- the
    product whose expansion src/gpu/cudaImageFunctions.cuh:216-252 writes
    out.  It maps moving-image coordinates to fixed-image coordinates; the
    search samples the moving image at its inverse."""
    th, sx, sy, sh, tx, ty, kx, ky = [float(v) for v in params]
    c, s = math.cos(th), math.sin(th)
    Tc = np.array([[1, 0, cx], [0, 1, cy], [0, 0, 1.0]])
    Ti = np.array([[1, 0, -cx], [0, 1, -cy], [0, 0, 1.0]])
    K = np.array([[1, 0, 0], [0, 1, 0], [kx, ky, 1.0]])
    T = np.array([[1, 0, tx], [0, 1, ty], [0, 0, 1.0]])
    R = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1.0]])
    Sh = np.array([[1, sh, 0], [0, 1, 0], [0, 0, 1.0]])
    Sc = np.diag([sx, sy, 1.0])
    H = Tc @ K @ T @ R @ Sh @ Sc @ Ti
    return H / H[2, 2]


def decompose(H, cx, cy):
    """Inverse of homography(): the 8 parameters of a homography (sx > 0,
    sy > 0).  With M = Tc^-1 H Tc = lambda K A (A affine): the keystone row
    solves [M31 M32] = kx [M11 M12] + ky [M21 M22], lambda = M33 - kx M13 -
    ky M23, and A = rows 1-2 of M / lambda = T R Sh Sc gives theta and sx
    from its first column, then sy and shear from R^-1 times its second."""
    Tc = np.array([[1, 0, cx], [0, 1, cy], [0, 0, 1.0]])
    Ti = np.array([[1, 0, -cx], [0, 1, -cy], [0, 0, 1.0]])
    M = Ti @ np.asarray(H, np.float64) @ Tc
    kx, ky = np.linalg.solve(np.array([[M[0, 0], M[1, 0]], [M[0, 1], M[1, 1]]]), [M[2, 0], M[2, 1]])
    lam = M[2, 2] - kx * M[0, 2] - ky * M[1, 2]
    A = M[:2] / lam
    sx = math.hypot(A[0, 0], A[1, 0])
    th = math.atan2(A[1, 0], A[0, 0])
    c, s = math.cos(th), math.sin(th)
    sy = -s * A[0, 1] + c * A[1, 1]
    sh = (c * A[0, 1] + s * A[1, 1]) / sy
    return np.array([th, sx, sy, sh, A[0, 2], A[1, 2], kx, ky])


def corners(w, h):
    return np.array([[0, 0], [w - 1, 0], [w - 1, h - 1], [0, h - 1]], np.float64)


def corner_offsets(H, w, h):
    """Where H moves the 4 image corners, minus the corners: (8,) as
    (dx0, dy0, dx1, dy1, ...), corners in the order of corners()."""
    q = corners(w, h)
    p = np.c_[q, np.ones(4)] @ np.asarray(H, np.float64).T
    return (p[:, :2] / p[:, 2:3] - q).reshape(8)


def offsets_to_params(off, w, h):
    q = corners(w, h)
    H = synth.dlt(q, q + np.asarray(off, np.float64).reshape(4, 2))
    return decompose(H, w / 2.0, h / 2.0)


def nearest_centre(params, start, res, count):
    """Per-axis sample index nearest to params, clamped to the grid."""
    i = np.rint((np.asarray(params, np.float64) - start) / np.where(res == 0, 1, res))
    return np.clip(i, 0, np.asarray(count) - 1).astype(np.int32)


def index_axes(index, count):
    """Grid index -> per-axis sample indices (axis 0 fastest)."""
    out, rem = [], int(index)
    for c in count:
        out.append(rem % int(c))
        rem //= int(c)
    return np.array(out, np.int32)


def label_offsets(indices, g, w=SIZE, h=SIZE):
    """Regression targets of grid indices: corner offsets of their
    homographies in units of OFFSET_SCALE px, float32 [n, 8]."""
    out = np.zeros((len(indices), 8), np.float32)
    for k, i in enumerate(indices):
        out[k] = corner_offsets(homography(g.getGridPoint(int(i)), w / 2.0, h / 2.0), w, h) / OFFSET_SCALE
    return out


# ------------------------------------------------------------------- data

def pair(k, heldout=False):
    """(moving, fixed) float32 256x256 of training (or held-out) pair k."""
    base = HELDOUT_SEED if heldout else TRAIN_SEED
    m, f, _ = synth.make_pair(base + int(k), SIZE, sigma=SIGMA, perturb=PERTURB, dtype=np.float32,
                              extra_shapes=True)
    return m, f


def _pair_job(args):
    k, heldout = args
    m, f = pair(k, heldout)
    return k, m, f


def grid():
    from .api import CudaGrid
    g = CudaGrid(8)
    g.setStartPoint(GRID_START)
    g.setEndPoint(GRID_END)
    g.setNumSamples(GRID_COUNT)
    return g


def full_grid_argmin(g, mov, fix, stream=None):
    """(value, index) of the plain full-grid NCC search, no cost tensor."""
    from . import _lib as L
    from .api import _stream_ptr, calcNCC, make_opts
    o = make_opts(calcNCC)
    out = L.Argmin()
    L.check(L.lib().cgs_search(ctypes.byref(g._g), ctypes.byref(mov._img), ctypes.byref(fix._img), ctypes.byref(o),
                               None, ctypes.byref(out), _stream_ptr(stream)))
    return out.value, out.index


def pair_stream(indices, heldout=False, workers=6):
    """Yields (k, moving, fixed) in order, generated by a spawn pool."""
    import multiprocessing as mp
    ctx = mp.get_context("spawn")
    with ctx.Pool(workers) as pool:
        for item in pool.imap(_pair_job, [(k, heldout) for k in indices], chunksize=4):
            yield item


def label(first, count, seconds, out, heldout=False, workers=6):
    """Full-grid argmins of pairs first..first+count-1 (stops early after
    `seconds`); writes {index, value, first, count} to out every 500 pairs."""
    from .api import CudaImage
    g = grid()
    idx = np.full(count, -1, np.int32)
    val = np.full(count, np.nan, np.float32)
    t0 = time.time()
    done = 0

    def save():
        tmp = out + ".tmp.npz"
        np.savez(tmp, index=idx[:done], value=val[:done], first=first, count=done, heldout=heldout,
                 generator=synth.GENERATOR_VERSION, grid_start=GRID_START, grid_end=GRID_END, grid_count=GRID_COUNT)
        os.replace(tmp, out)

    for k, m, f in pair_stream(range(first, first + count), heldout, workers):
        v, i = full_grid_argmin(g, CudaImage.from_numpy(m), CudaImage.from_numpy(f))
        idx[k - first], val[k - first] = i, v
        done = k - first + 1
        if done % 500 == 0:
            save()
            el = time.time() - t0
            print(f"[label] {done} pairs, {el:.0f} s, {done / el:.2f} pairs/s", flush=True)
        if time.time() - t0 > seconds:
            break
    save()
    print(f"[label] finished {done} pairs in {time.time() - t0:.0f} s -> {out}", flush=True)
    return done


# ------------------------------------------------------ inputs, training

def load_labels(paths):
    """Concatenated (pair numbers, argmin indices, argmin costs) of label files."""
    ks, idx, val = [], [], []
    for p in paths:
        with np.load(p, allow_pickle=False) as z:
            n = int(z["count"])
            ks.append(int(z["first"]) + np.arange(n))
            idx.append(z["index"][:n])
            val.append(z["value"][:n])
    return np.concatenate(ks), np.concatenate(idx).astype(np.int32), np.concatenate(val)


def upload_pairs(ks, heldout=False, workers=12, keep_images=False, log=None):
    """Generates pairs ks (CPU pool), uploads them and builds the network
    inputs on the GPU.  Returns (X int16 [n, 128, 128, 2], images) where
    images is a list of (moving, fixed) CudaImage when keep_images."""
    import torch
    from .api import CudaImage
    from .cnn import ProposalNet, SIDE
    n = len(ks)
    X = torch.empty((n, SIDE, SIDE, 2), dtype=torch.int16, device="cuda")
    images = [] if keep_images else None
    chunk_m, chunk_f, start = [], [], 0
    t0 = time.time()
    for j, (k, m, f) in enumerate(pair_stream(ks, heldout, workers)):
        cm, cf = CudaImage.from_numpy(m), CudaImage.from_numpy(f)
        chunk_m.append(cm)
        chunk_f.append(cf)
        if keep_images:
            images.append((cm, cf))
        if len(chunk_m) == 256 or j == n - 1:
            ProposalNet.prepare(chunk_m, chunk_f, out=X[start:start + len(chunk_m)])
            start += len(chunk_m)
            chunk_m, chunk_f = [], []
            if log and (start % 4096 < 256 or start == n):
                log(f"[data] {start}/{n} pairs ready, {time.time() - t0:.0f} s")
    return X, images


def train(net, X, Y, epochs, batch=128, lr=1e-3, lr_min=5e-5, seed=0, seconds=None, Xval=None, Yval=None,
          log=print):
    """Adam on the mean squared corner-offset error, cosine learning rate
    from lr to lr_min; keeps the parameters with the lowest validation loss
    when a validation set is given.  X: int16 [n, 128, 128, 2]; Y: float32
    [n, 8] (device).  Returns a history of (step, train loss, val loss)."""
    import torch
    n = int(X.shape[0])
    steps_per_epoch = max(1, n // batch)
    total = epochs * steps_per_epoch
    gen = torch.Generator(device="cpu")
    gen.manual_seed(seed)
    hist, best, best_p = [], float("inf"), None
    t0 = time.time()
    step = 0
    for ep in range(epochs):
        perm = torch.randperm(n, generator=gen).to(X.device)
        run = 0.0
        for b in range(steps_per_epoch):
            idx = perm[b * batch:(b + 1) * batch]
            net.forward(X[idx].contiguous())
            want = b == steps_per_epoch - 1
            loss = net.backward(Y[idx].contiguous(), want_loss=want)
            rate = lr_min + 0.5 * (lr - lr_min) * (1 + math.cos(math.pi * step / max(1, total)))
            net.adam_step(rate)
            step += 1
            if want:
                run = loss
        vl = evaluate_loss(net, Xval, Yval, batch) if Xval is not None else float("nan")
        hist.append((step, run, vl))
        if Xval is not None and vl < best:
            best, best_p = vl, net.get_params()
        log(f"[train] epoch {ep + 1}/{epochs} step {step} loss {run:.5f} val {vl:.5f} {time.time() - t0:.0f} s")
        if seconds is not None and time.time() - t0 > seconds:
            break
    if best_p is not None:
        net.set_params(best_p)
    return hist


def evaluate_loss(net, X, Y, batch=256):
    import torch
    tot, n = 0.0, int(X.shape[0])
    for b in range(0, n, batch):
        out = net.forward(X[b:b + batch].contiguous())
        tot += float(((out - Y[b:b + batch]) ** 2).sum())
    torch.cuda.synchronize()
    return tot / (8 * n)


# --------------------------------------------------------------- pipeline

def predict_offsets(net, X, batch=256):
    """Corner offsets in px, numpy [n, 8]."""
    import torch
    outs = []
    for b in range(0, int(X.shape[0]), batch):
        outs.append(net.forward(X[b:b + batch].contiguous()))
    return torch.cat(outs).cpu().numpy().astype(np.float64) * OFFSET_SCALE


def centres_from_offsets(off, g, w=SIZE, h=SIZE):
    start, res, count = g.start_point(), g.resolution(), g.getAxisSampleCounts()
    return np.stack([nearest_centre(offsets_to_params(o, w, h), start, res, count) for o in off])


def lattice_coords(off, g, w=SIZE, h=SIZE):
    """Continuous per-axis sample coordinates (p - start) / res of the
    proposals, [n, dim]."""
    start, res = g.start_point().astype(np.float64), g.resolution().astype(np.float64)
    return np.stack([(offsets_to_params(o, w, h) - start) / np.where(res == 0, 1, res) for o in off])


START_MODES = ("nearest", "opposite", "identity", "identity_far")


def make_centres(off, g, modes=("nearest", "opposite", "identity"), w=SIZE, h=SIZE):
    """Descent starts per pair, [n, len(modes), dim]: "nearest" = the grid
    point nearest the proposal; "opposite" = the far corner of the lattice
    cell holding the proposal (floor and ceil swapped on every axis);
    "identity" = the grid point nearest the identity homography;
    "identity_far" = the far corner of the lattice cell holding the identity."""
    count = g.getAxisSampleCounts().astype(np.int64)
    u = lattice_coords(off, g, w, h)
    near = np.clip(np.rint(u), 0, count - 1).astype(np.int32)
    out = []
    for m in modes:
        if m == "nearest":
            out.append(near)
        elif m == "opposite":
            other = np.where(np.rint(u) > u, np.floor(u), np.ceil(u))
            out.append(np.clip(other, 0, count - 1).astype(np.int32))
        elif m == "identity":
            out.append(identity_centres(len(off), g))
        elif m == "identity_far":
            out.append(fixed_centres(len(off), g, 2)[:, 1, :])
        else:
            raise ValueError(f"unknown start mode {m!r}")
    return np.stack(out, axis=1)


def identity_centres(n, g):
    """Centres without a network: the grid point nearest the identity."""
    c = nearest_centre([0, 1, 1, 0, 0, 0, 0, 0], g.start_point(), g.resolution(), g.getAxisSampleCounts())
    return np.tile(c, (n, 1))


def refine(images, g, centres, threshold=float("inf"), max_steps=8, fallback=True):
    """Batched refinement of all pairs; centres [n, dim] or [n, starts, dim]."""
    from .api import BatchSearcher, calcNCC
    b = BatchSearcher()
    return b.refine([(calcNCC, g, m, f, None) for m, f in images], centres, radius=1, max_steps=max_steps,
                    threshold=threshold, fallback=fallback)


def calibrate_threshold(values, fallback_rate):
    """Accepted-cost threshold that sends the worst-matching fallback_rate of
    pairs to the full grid: the (1 - fallback_rate) quantile of the local
    minima's costs (lower is better), taken over validation pairs."""
    v = np.sort(np.asarray(values, np.float64))
    if len(v) == 0 or fallback_rate <= 0:
        return float("inf")
    k = int(math.ceil((1.0 - fallback_rate) * len(v))) - 1
    return float(v[min(max(k, 0), len(v) - 1)])


def run_learned(net, images, g, threshold, max_steps=8, batch=256, modes=("nearest",)):
    """The learned pipeline on resident pairs: inputs, network, descent
    starts (make_centres modes), local refinement with full-grid fallback.
    Returns (results, seconds, offsets)."""
    import torch
    from .cnn import ProposalNet
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    X = torch.empty((len(images), 128, 128, 2), dtype=torch.int16, device="cuda")
    for b in range(0, len(images), batch):
        part = images[b:b + batch]
        ProposalNet.prepare([m for m, _ in part], [f for _, f in part], out=X[b:b + len(part)])
    off = predict_offsets(net, X, batch)
    res = refine(images, g, make_centres(off, g, modes), threshold, max_steps)
    torch.cuda.synchronize()
    return res, time.perf_counter() - t0, off


def fixed_centres(n, g, starts):
    """Data-independent starts: the corners of the lattice cell holding the
    identity homography -- its nearest grid point, the opposite corner, then
    corners alternating per axis -- [n, starts, dim]."""
    count = g.getAxisSampleCounts().astype(np.int64)
    start, res = g.start_point().astype(np.float64), g.resolution().astype(np.float64)
    u = (np.array([0, 1, 1, 0, 0, 0, 0, 0], np.float64)[:len(count)] - start) / np.where(res == 0, 1, res)
    near, far = np.rint(u), np.where(np.rint(u) > u, np.floor(u), np.ceil(u))
    cands = [near, far]
    for k in range(starts):
        mask = ((np.arange(len(count)) >> k) & 1).astype(bool)
        cands.append(np.where(mask, far, near))
    out = [np.clip(c, 0, count - 1).astype(np.int32) for c in cands[:starts]]
    return np.tile(np.stack(out)[None], (n, 1, 1))


def run_without_network(images, g, threshold, max_steps=8, starts=1):
    """Ablation: the same refinement and fallback from as many fixed starts
    (fixed_centres) instead of the network's."""
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = refine(images, g, fixed_centres(len(images), g, starts), threshold, max_steps)
    torch.cuda.synchronize()
    return res, time.perf_counter() - t0


def run_full_grid(images, g):
    """The arbiter: plain full-grid argmins, one search per pair."""
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = np.array([full_grid_argmin(g, m, f)[1] for m, f in images], np.int32)
    torch.cuda.synchronize()
    return out, time.perf_counter() - t0


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m cugridsearch_amd.proposal")
    sub = ap.add_subparsers(dest="cmd", required=True)
    lb = sub.add_parser("label", help="full-grid NCC argmins of training pairs (GPU)")
    lb.add_argument("--first", type=int, default=0)
    lb.add_argument("--count", type=int, default=1000)
    lb.add_argument("--seconds", type=float, default=1e9)
    lb.add_argument("--heldout", action="store_true")
    lb.add_argument("--workers", type=int, default=6)
    lb.add_argument("--out", required=True)
    tr = sub.add_parser("train", help="train the network on labelled pairs (GPU)")
    tr.add_argument("--labels", nargs="+", required=True)
    tr.add_argument("--epochs", type=int, default=40)
    tr.add_argument("--seconds", type=float, default=None)
    tr.add_argument("--batch", type=int, default=128)
    tr.add_argument("--lr", type=float, default=1e-3)
    tr.add_argument("--val", type=int, default=1024, help="labelled pairs held back for validation / threshold")
    tr.add_argument("--fallback-rate", type=float, default=0.01,
                    help="share of validation pairs whose local minimum the threshold sends to the full grid")
    tr.add_argument("--modes", default="nearest,opposite", help="descent starts (make_centres modes)")
    tr.add_argument("--seed", type=int, default=1)
    tr.add_argument("--workers", type=int, default=12)
    tr.add_argument("--out", required=True)
    ev = sub.add_parser("evaluate", help="learned pipeline vs full grid on held-out pairs (GPU)")
    ev.add_argument("--weights", required=True)
    ev.add_argument("--pairs", type=int, default=1024)
    ev.add_argument("--max-steps", type=int, default=8)
    ev.add_argument("--threshold", type=float, default=None, help="default: the one saved with the weights")
    ev.add_argument("--modes", default=None, help="descent starts; default: the ones saved with the weights")
    ev.add_argument("--workers", type=int, default=12)
    ev.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    if a.cmd == "label":
        os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
        label(a.first, a.count, a.seconds, a.out, a.heldout, a.workers)
    elif a.cmd == "train":
        train_main(a)
    elif a.cmd == "evaluate":
        evaluate_main(a)
    return 0


def evaluate(net, images, g, threshold, max_steps=8, reference_argmin=None, log=None, modes=("nearest",)):
    """Agreement with the full-grid argmin and pairs/s of: the plain full
    grid (the arbiter), the learned pipeline, the same refinement from as many
    fixed starts (no network), and the network's nearest grid point alone."""
    n = len(images)
    if reference_argmin is None:
        reference_argmin, t_full = run_full_grid(images, g)
    else:
        t_full = None
    run_learned(net, images[:min(8, n)], g, threshold, max_steps, modes=modes)   # first-call set-up untimed
    res, t_learn, off = run_learned(net, images, g, threshold, max_steps, modes=modes)
    abl, t_abl = run_without_network(images, g, threshold, max_steps, starts=len(modes))
    count = g.getAxisSampleCounts().astype(int)
    inc = np.cumprod(np.r_[1, count[:-1]])
    cnn_idx = (centres_from_offsets(off, g) * inc).sum(1)
    full_pts = g.numElements()

    def summary(r, t):
        return dict(agree=float((r["index"] == reference_argmin).mean()), seconds=t, pairs_per_s=n / t,
                    local=float((r["outcome"] == 0).mean()), fallback_boundary=float((r["outcome"] == 1).mean()),
                    fallback_threshold=float((r["outcome"] == 2).mean()), mean_steps=float(r["steps"].mean()),
                    points_per_pair=float(r["points"].mean()), points_vs_full=float(r["points"].mean() / full_pts))

    rep = dict(pairs=n, grid_points=full_pts, threshold=float(threshold), max_steps=max_steps, starts=list(modes),
               learned=summary(res, t_learn), without_network=summary(abl, t_abl),
               network_alone_agree=float((cnn_idx == reference_argmin).mean()))
    if t_full is not None:
        rep["full_grid"] = dict(seconds=t_full, pairs_per_s=n / t_full)
        rep["speedup_learned_vs_full"] = t_full / t_learn
        rep["speedup_without_network_vs_full"] = t_full / t_abl
    if log:
        log("[evaluate] " + str(rep))
    return rep, res, reference_argmin


def saved_settings(path):
    """(threshold, descent-start modes) stored with trained weights."""
    with np.load(path, allow_pickle=False) as z:
        thr = float(z["threshold"]) if "threshold" in z else float("inf")
        modes = tuple(str(z["modes"]).split(",")) if "modes" in z else ("nearest",)
    return thr, modes


def evaluate_main(a):
    import json
    from .cnn import ProposalNet
    log = lambda m: print(m, flush=True)  # noqa: E731
    net = ProposalNet.load(a.weights, max_batch=256)
    thr, modes = saved_settings(a.weights)
    if a.threshold is not None:
        thr = a.threshold
    if a.modes is not None:
        modes = tuple(a.modes.split(","))
    g = grid()
    _, images = upload_pairs(range(a.pairs), heldout=True, workers=a.workers, keep_images=True, log=log)
    rep, _, _ = evaluate(net, images, g, thr, a.max_steps, log=log, modes=modes)
    if a.out:
        with open(a.out, "w") as f:
            json.dump(rep, f, indent=1)


def train_main(a):
    import json
    import torch
    from .cnn import ProposalNet
    log = lambda m: print(m, flush=True)  # noqa: E731
    ks, idx, _ = load_labels(a.labels)
    g = grid()
    ok = idx >= 0
    ks, idx = ks[ok], idx[ok]
    nval = min(a.val, len(ks) // 5)
    log(f"[train] {len(ks)} labelled pairs, {nval} for validation")
    X, images = upload_pairs(ks, workers=a.workers, keep_images=False, log=log)
    Y = torch.from_numpy(label_offsets(idx, g)).cuda()
    Xtr, Ytr, Xval, Yval = X[nval:], Y[nval:], X[:nval], Y[:nval]
    net = ProposalNet(max_batch=max(256, a.batch), seed=a.seed)
    hist = train(net, Xtr, Ytr, a.epochs, a.batch, a.lr, seed=a.seed, seconds=a.seconds, Xval=Xval, Yval=Yval,
                 log=log)
    # threshold on the validation pairs: local minima whose index differs
    # from the full-grid label are the errors a threshold has to catch
    del X, Xtr, Ytr
    _, vimages = upload_pairs(ks[:nval], workers=a.workers, keep_images=True)
    modes = tuple(a.modes.split(","))
    vres, vsec, _ = run_learned(net, vimages, g, threshold=float("inf"), modes=modes)
    local = vres["outcome"] == 0
    wrong = vres["local_index"] != idx[:nval]
    thr = calibrate_threshold(vres["local_value"][local], a.fallback_rate)
    accepted = local & (vres["local_value"] <= thr)
    meta = dict(train_pairs=int(len(ks) - nval), val_pairs=int(nval), epochs_run=len(hist),
                history=[list(map(float, h)) for h in hist], threshold=thr, fallback_rate=a.fallback_rate,
                val_local_rate=float(local.mean()), val_local_wrong=float((local & wrong).mean()),
                val_agree_inf=float((vres["index"] == idx[:nval]).mean()),
                val_agree_threshold=float(((accepted & ~wrong) | ~accepted).mean()),
                labels=[os.path.basename(p) for p in a.labels], modes=list(modes), val_seconds=vsec)
    log("[train] " + json.dumps({k: v for k, v in meta.items() if k != "history"}))
    net.save(a.out, threshold=np.float32(thr), modes=",".join(modes), meta=json.dumps(meta))
    log(f"[train] saved {a.out}")


if __name__ == "__main__":
    sys.exit(main())
