"""Seeded synthetic image pairs and grids for the search workloads.

Recipe (all float64 until the final cast):
  fixed  = U[0,1) noise -> separable Gaussian blur (radius ceil(3 sigma),
           reflect-101 borders) -> min-max rescale to [0, 255]
           (-> round -> uint8 for uint8 pairs);
  moving = fixed sampled through H_true (bilinear, 0 outside), where H_true is
           the 4-point DLT homography taking the image corners to the corners
           perturbed uniformly by +-perturb px; moving(q) = fixed(H_true q), so
           the search's inverse warp of the fixed grid recovers H_true.
  The mutual-information pair also gets m' = 255 (m/255)^gamma,
  gamma ~ U(0.5, 2.0), then + N(0, noise^2), clipped and rounded.

Generator version is recorded in every fixture manifest (GENERATOR_VERSION).
"""
import math

import numpy as np

GENERATOR_VERSION = 1

# Parameter order of the 8-D grid: theta, scaleX, scaleY, shearXY,
# translateX, translateY, keystoneX, keystoneY (reference
# src/gpu/cudaImageFunctions.cuh:222-229).
DEG5 = 0.0872665


def grid_ranges(levels, translate):
    """Start/end/count rows for the homography grid: theta +-5 deg, scale
    0.95..1.05, shear +-0.05, translation +-translate px, keystone +-1e-3."""
    start = np.array([-DEG5, 0.95, 0.95, -0.05, -translate, -translate, -1e-3, -1e-3], np.float32)
    end = np.array([DEG5, 1.05, 1.05, 0.05, translate, translate, 1e-3, 1e-3], np.float32)
    count = np.full(8, levels, np.float32)
    return start, end, count


def gaussian_blur(img, sigma):
    r = int(math.ceil(3 * sigma))
    t = np.arange(-r, r + 1, dtype=np.float64)
    k = np.exp(-0.5 * (t / sigma) ** 2)
    k /= k.sum()
    out = img
    for axis in (0, 1):
        pad = [(0, 0), (0, 0)]
        pad[axis] = (r, r)
        p = np.pad(out, pad, mode="reflect")  # reflect-101
        acc = np.zeros_like(out)
        n = out.shape[axis]
        for i, kv in enumerate(k):
            sl = [slice(None), slice(None)]
            sl[axis] = slice(i, i + n)
            acc += kv * p[tuple(sl)]
        out = acc
    return out


def dlt(src, dst):
    """FP64 4-point DLT: H with dst ~ H src (h33 = 1)."""
    A, b = [], []
    for (x, y), (u, v) in zip(src, dst):
        A.append([x, y, 1, 0, 0, 0, -u * x, -u * y])
        b.append(u)
        A.append([0, 0, 0, x, y, 1, -v * x, -v * y])
        b.append(v)
    h = np.linalg.solve(np.array(A, np.float64), np.array(b, np.float64))
    return np.append(h, 1.0).reshape(3, 3)


def bilinear_zero(img, xs, ys):
    h, w = img.shape
    x0 = np.floor(xs).astype(np.int64)
    y0 = np.floor(ys).astype(np.int64)
    ax = xs - x0
    ay = ys - y0
    out = np.zeros(xs.shape, np.float64)
    for dy, wy in ((0, 1 - ay), (1, ay)):
        for dx, wx in ((0, 1 - ax), (1, ax)):
            xi, yi = x0 + dx, y0 + dy
            ok = (xi >= 0) & (xi < w) & (yi >= 0) & (yi < h)
            v = np.zeros(xs.shape, np.float64)
            v[ok] = img[yi[ok], xi[ok]]
            out += wx * wy * v
    inside = (xs >= 0) & (ys >= 0) & (xs <= w - 1) & (ys <= h - 1)
    return np.where(inside, out, 0.0)


def warp_by(img, H, out_shape):
    hh, ww = out_shape
    ys, xs = np.mgrid[0:hh, 0:ww].astype(np.float64)
    den = H[2, 0] * xs + H[2, 1] * ys + H[2, 2]
    u = (H[0, 0] * xs + H[0, 1] * ys + H[0, 2]) / den
    v = (H[1, 0] * xs + H[1, 1] * ys + H[1, 2]) / den
    return bilinear_zero(img, u, v)


def make_pair(seed, width, height=None, sigma=3.0, perturb=3.0, dtype=np.uint8, gamma=False, noise=0.0,
              extra_shapes=False):
    """Returns (moving, fixed, H_true) as (H x W) arrays of dtype."""
    height = width if height is None else height
    rng = np.random.default_rng(seed)
    f = gaussian_blur(rng.random((height, width)), sigma)
    if extra_shapes:
        for _ in range(int(rng.integers(2, 6))):
            x0, x1 = np.sort(rng.integers(0, width, 2))
            y0, y1 = np.sort(rng.integers(0, height, 2))
            f[y0:y1 + 1, x0:x1 + 1] += rng.uniform(-1.0, 1.0) * f.std() * 2.0
        for _ in range(int(rng.integers(1, 4))):
            a, c = rng.uniform(-1, 1, 2)
            ys, xs = np.mgrid[0:height, 0:width]
            f += (a * (xs - width / 2) + c * (ys - height / 2) > 0) * rng.uniform(-1, 1) * f.std()
    f = (f - f.min()) / max(f.max() - f.min(), 1e-12) * 255.0
    corners = np.array([[0, 0], [width - 1, 0], [width - 1, height - 1], [0, height - 1]], np.float64)
    moved_c = corners + rng.uniform(-perturb, perturb, corners.shape)
    H_true = dlt(corners, moved_c)
    m = warp_by(f, H_true, (height, width))
    if gamma:
        g = rng.uniform(0.5, 2.0)
        m = 255.0 * (m / 255.0) ** g
    if noise > 0:
        m = m + rng.normal(0.0, noise, m.shape)
    if dtype == np.uint8:
        f = np.clip(np.rint(f), 0, 255).astype(np.uint8)
        m = np.clip(np.rint(m), 0, 255).astype(np.uint8)
    else:
        f = f.astype(np.float32)
        m = np.clip(m, 0.0, 255.0).astype(np.float32)
    return m, f, H_true


# The three single-pair workloads.  Seeds are fixed so that fixtures, tests
# and bench runs see identical inputs.
WORKLOADS = {
    "nn_sad_64": dict(seed=0xC6501, width=64, sigma=3.0, perturb=3.0, dtype=np.uint8,
                      levels=4, translate=4.0, cost="sad", sampler="nearest"),
    "bilinear_ncc_256": dict(seed=0xC6502, width=256, sigma=6.0, perturb=8.0, dtype=np.float32,
                             levels=6, translate=8.0, cost="ncc", sampler="bilinear"),
    "bilinear_mi32_512": dict(seed=0xC6503, width=512, sigma=6.0, perturb=8.0, dtype=np.uint8,
                              levels=5, translate=8.0, cost="mi", sampler="bilinear", bins=32,
                              gamma=True, noise=4.0),
}


def workload(name, levels=None, width=None):
    """(moving, fixed, start, end, count, spec) for a named workload; levels
    and width may be overridden for reduced-size parity cases."""
    spec = dict(WORKLOADS[name])
    if levels is not None:
        spec["levels"] = levels
    if width is not None:
        spec["width"] = width
    m, f, _ = make_pair(spec["seed"], spec["width"], sigma=spec["sigma"], perturb=spec["perturb"],
                        dtype=spec["dtype"], gamma=spec.get("gamma", False), noise=spec.get("noise", 0.0))
    start, end, count = grid_ranges(spec["levels"], spec["translate"])
    return m, f, start, end, count, spec


def batch_workload(n=256, seed=0xC6504, sides=(64, 512), levels=(4, 5, 6)):
    """Ragged batch: n pairs with width and height drawn independently from
    U{sides}, uint8 and float32 alternating, grid levels from `levels`, costs
    SAD / NCC / MI 1:1:1 in a seeded shuffle.  Returns a list of dicts with
    moving, fixed, start, end, count, cost ("sad" | "ncc" | "mi")."""
    rng = np.random.default_rng(seed)
    costs = (["sad", "ncc", "mi"] * ((n + 2) // 3))[:n]
    rng.shuffle(costs)
    out = []
    for k in range(n):
        w = int(rng.integers(sides[0], sides[1] + 1))
        h = int(rng.integers(sides[0], sides[1] + 1))
        dtype = np.uint8 if k % 2 == 0 else np.float32
        lv = int(rng.choice(levels))
        sigma = 3.0 if max(w, h) <= 128 else 6.0
        perturb = max(2.0, min(w, h) / 32.0)
        m, f, _ = make_pair(int(rng.integers(1 << 31)), w, h, sigma=sigma, perturb=perturb, dtype=dtype,
                            gamma=costs[k] == "mi", noise=4.0 if costs[k] == "mi" else 0.0)
        start, end, count = grid_ranges(lv, max(2.0, min(w, h) / 32.0))
        out.append(dict(moving=m, fixed=f, start=start, end=end, count=count, cost=costs[k], levels=lv))
    return out
