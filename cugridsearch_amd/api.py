"""Host-side mirror of the reference's search API, over the C-ABI.

Names, argument meaning and call order follow the reference so that a driver
reads the same (compare src/gpu/cudaGridSearch_main_mi.cu:300-345):

    grid = CudaGrid(8)
    grid.setStartPoint(start); grid.setEndPoint(end); grid.setNumSamples(counts)
    func_values = CudaTensor(grid.getAxisSampleCounts())
    searcher = CudaGridSearcher(grid, func_values)
    searcher.search_by_value(calcNCC, image_mov, image_fix)
    value, index = func_values.find_extrema()
    point = grid.getGridPoint(index)

Device memory is held in torch tensors (allocation and streams only); all
compute runs in libcgs.so.  Functors are tags naming a cost of the C-ABI
(the reference binds a __device__ function pointer instead,
src/gpu/cudaGridSearch.cuh:284-285).
"""
import ctypes
import os
import dataclasses

import numpy as np

from . import _lib as L


def _torch():
    import torch
    return torch


def _stream_ptr(stream):
    if stream is None:
        return None
    return ctypes.c_void_p(stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream))


# ----------------------------------------------------------------- functors

@dataclasses.dataclass(frozen=True)
class Functor:
    """A cost functor: which cost, how the grid point becomes a warp, and
    which CudaImage sampler reads the moving image."""
    name: str
    cost: int
    param_mode: int
    sampler: int = L.SAMPLER_BILINEAR
    bins: int = 64

    def with_(self, **kw):
        return dataclasses.replace(self, **kw)


# src/gpu/cudaErrorFunctions.cuh:31-87 (non-centred H, valueAt = bilinear)
averageAbsoluteDifference = Functor("averageAbsoluteDifference", L.COST_SAD, L.PARAM_NONCENTRED)
# src/gpu/cudaErrorFunctions.cuh:194-224 (raw forward H)
averageAbsoluteDifferenceH = Functor("averageAbsoluteDifferenceH", L.COST_SAD, L.PARAM_RAW_H)
# src/gpu/cudaErrorFunctions.cuh:89-114 (2-D translation)
sumOfAbsoluteDifferences = Functor("sumOfAbsoluteDifferences", L.COST_SAD_SUM, L.PARAM_TRANSLATION)
# src/gpu/cudaErrorFunctions.cuh:378-446
calcSQD = Functor("calcSQD", L.COST_SQD, L.PARAM_CENTRED)
# src/gpu/cudaErrorFunctions.cuh:448-507
calcNCC = Functor("calcNCC", L.COST_NCC, L.PARAM_CENTRED)
# src/gpu/cudaErrorFunction_miStreams.cuh:212-368 (64 bins)
calcMIstream = Functor("calcMIstream", L.COST_MI, L.PARAM_CENTRED, bins=64)
# src/gpu/cudaErrorFunction_mi.cuh:8-115
calcMI = Functor("calcMI", L.COST_MI_PLAIN, L.PARAM_CENTRED, bins=64)
# the thread-per-row "stream" functors: the same sums, their own final combine
# src/gpu/cudaErrorFunctionsStreams.cuh:31-83 (2-D translation; no overlap -> 0)
averageAbsoluteDifference_stream = Functor("averageAbsoluteDifference_stream", L.COST_SAD_STREAM, L.PARAM_TRANSLATION)
# src/gpu/cudaErrorFunctionsStreams.cuh:85-163 (0 unless both sums of squares are positive)
calcNCCstream = Functor("calcNCCstream", L.COST_NCC_STREAM, L.PARAM_CENTRED)
# src/gpu/cudaErrorFunctionsStreams.cuh:283-351 (no overlap sentinel: 0/0)
calcSQDstream = Functor("calcSQDstream", L.COST_SQD_STREAM, L.PARAM_CENTRED)
# line traversal: each fixed row walks its line through the moving image in
# unit lambda steps, with an overlap test and a sentinel
# src/gpu/cudaErrorFunctions.cuh:226-300 (123123)
calcSQDAlt = Functor("calcSQDAlt", L.COST_SQD_ALT, L.PARAM_CENTRED)
# src/gpu/cudaErrorFunctions.cuh:302-376 (123123)
calcNCCAlt = Functor("calcNCCAlt", L.COST_NCC_ALT, L.PARAM_CENTRED)
# src/gpu/cudaErrorFunctionsStreams.cuh:165-281 (0)
calcNCCstreamAlt = Functor("calcNCCstreamAlt", L.COST_NCC_STREAM_ALT, L.PARAM_CENTRED)
# src/gpu/cudaErrorFunction_miStreams.cuh:9-210 (64 bins, 123123)
calcMIstreamAlt = Functor("calcMIstreamAlt", L.COST_MI_STREAM_ALT, L.PARAM_CENTRED, bins=64)
# src/gpu/cudaErrorFunction_mi.cuh:117-255: normalised H and lines, -MI * fixed area (64 bins)
calcMIAlt = Functor("calcMIAlt", L.COST_MI_ALT, L.PARAM_NORM, bins=64)
# the centred-homography SAD used by the search workloads
centredAbsoluteDifference = Functor("centredAbsoluteDifference", L.COST_SAD, L.PARAM_CENTRED)


# --------------------------------------------------------------- containers

class CudaGrid:
    """CudaGrid<float, D> (src/gpu/cudaGridSearch.cuh:74-231)."""

    def __init__(self, dim=8):
        if not 1 <= dim <= L.MAX_DIM:
            raise ValueError("grid dimension must be 1..8")
        self._dim = dim
        self._start = [0.0] * dim
        self._end = [0.0] * dim
        self._count = [1.0] * dim
        self._g = L.Grid()
        self.update()

    def update(self):
        L.check(L.lib().cgs_grid_init(ctypes.byref(self._g), self._dim, L.floats(self._start),
                                      L.floats(self._end), L.floats(self._count)))

    def setStartPoint(self, v):
        self._start = [float(x) for x in v]
        self.update()

    def setEndPoint(self, v):
        self._end = [float(x) for x in v]
        self.update()

    def setNumSamples(self, v):
        self._count = [float(x) for x in v]
        self.update()

    def getDimension(self):
        return self._dim

    def numElements(self):
        return int(self._g.num_elements)

    def getAxisSampleCounts(self):
        return np.array(self._g.count[:self._dim], np.float32)

    def start_point(self):
        return np.array(self._g.start[:self._dim], np.float32)

    def end_point(self):
        return np.array(self._g.end[:self._dim], np.float32)

    def resolution(self):
        return np.array(self._g.resolution[:self._dim], np.float32)

    def getGridPoint(self, index):
        out = (ctypes.c_float * self._dim)()
        L.check(L.lib().cgs_grid_point(ctypes.byref(self._g), int(index), out))
        return np.array(out[:], np.float32)

    def subbox(self, lo, count):
        """The grid's samples lo[a] .. lo[a]+count[a]-1 on every axis as a grid
        of its own (cgs_grid_subbox): same points, bit for bit."""
        g = CudaGrid.__new__(CudaGrid)
        g._dim = self._dim
        g._start, g._end, g._count = self._start, self._end, self._count
        g._g = L.Grid()
        lo_a = (ctypes.c_int32 * self._dim)(*[int(v) for v in lo])
        cnt_a = (ctypes.c_int32 * self._dim)(*[int(v) for v in count])
        L.check(L.lib().cgs_grid_subbox(ctypes.byref(self._g), lo_a, cnt_a, ctypes.byref(g._g)))
        return g

    def offset(self):
        return np.array(self._g.offset[:self._dim], np.int32)

    def display(self, name="", echo=True):
        """CudaMatrix::display format (include/cudaTensor.cuh:395-415): the four
        rows start / end / resolution / count.  Prints it (echo) and returns it."""
        rows = [self.start_point(), self.end_point(), self.resolution(), self.getAxisSampleCounts()]
        lines = [f"Matrix {name} 4 x {self._dim} elements of f", ""]
        for r in rows:
            lines.append("{ " + ", ".join(f"{v:g}" for v in r) + " }")
        text = "\n".join(lines) + "\n"
        if echo:
            print(text)
        return text


class CudaTensor:
    """CudaTensor<float, D> (include/cudaTensor.cuh:108-361) built from the
    grid's axis counts: element order = grid index order (axis 0 fastest)."""

    def __init__(self, dims, device="cuda", data=None):
        torch = _torch()
        self.dims = [int(d) for d in dims]
        n = 1
        for d in self.dims:
            n *= d
        self._total = n
        self.data = data if data is not None else torch.empty(n, dtype=torch.float32, device=device)
        self.last_argmin = None

    def size(self, dim=-1):
        return self._total if dim == -1 else self.dims[dim]

    def bytesSize(self):
        return self._total * 4

    def find_extrema(self, tie_mode=L.TIE_LOWEST, stream=None):
        """Returns (min value, int32 index) -- include/cudaTensor.cuh:215-265."""
        out = L.Argmin()
        L.check(L.lib().cgs_find_extrema(ctypes.c_void_p(self.data.data_ptr()), self._total, tie_mode,
                                         ctypes.byref(out), _stream_ptr(stream)))
        return out.value, out.index


class CudaImage:
    """CudaImage<pix, C>(H, W) (include/cudaTensor.cuh:507-596) over a device
    tensor of shape (H, W) or (H, W, C), uint8 or float32."""

    def __init__(self, tensor):
        torch = _torch()
        if tensor.dtype not in (torch.uint8, torch.float32):
            raise TypeError("CudaImage holds uint8 or float32 pixels")
        self.tensor = tensor.contiguous()
        h, w = self.tensor.shape[:2]
        c = self.tensor.shape[2] if self.tensor.dim() == 3 else 1
        self._img = L.Image(self.tensor.data_ptr(), w, h, c,
                            L.PIX_U8 if tensor.dtype == torch.uint8 else L.PIX_F32)

    @classmethod
    def from_numpy(cls, arr, device="cuda"):
        return cls(_torch().from_numpy(np.ascontiguousarray(arr)).to(device))

    def width(self):
        return self._img.width

    def height(self):
        return self._img.height

    def channels(self):
        return self._img.channels


def filter_image(img, filt, out=None, stream=None):
    """CudaImage::filter with every channel filtered (include/cudaTensor.cuh:577-593,
    filterProcess include/cudaTensorKernels.cuh:353-392).  filt: (fh, fw)
    array of at most 256 taps.  Returns out (a new CudaImage when None)."""
    f = np.ascontiguousarray(filt, np.float32)
    if f.ndim != 2:
        raise ValueError("filter must be a (rows, cols) array")
    if out is None:
        out = CudaImage(_torch().empty_like(img.tensor))
    L.check(L.lib().cgs_image_filter(ctypes.byref(img._img), ctypes.byref(out._img), L.floats(f.reshape(-1)),
                                     f.shape[1], f.shape[0], _stream_ptr(stream)))
    return out


def resize_image(img, width, height, stream=None):
    """stbir_resize_uint8 (stb_image_resize v0.97, as the matcher calls it at
    src/cpu/multispectral_ImageMatcher.cu:531 and :550): a new uint8 CudaImage
    of width x height, byte for byte the library's resampling."""
    t = img.tensor
    shape = (int(height), int(width)) + tuple(t.shape[2:])
    out = CudaImage(_torch().empty(shape, dtype=t.dtype, device=t.device))
    L.check(L.lib().cgs_image_resize(ctypes.byref(img._img), ctypes.byref(out._img), _stream_ptr(stream)))
    return out


def warp_image(img, H, width, height, out=None, stream=None):
    """transformImage (src/gpu/cudaImageFunctions.cuh:40-62): img under the
    8-entry homography H into a width x height image (a new zero-filled one
    when out is None, as writeTransformedImageToDisk starts from)."""
    if out is None:
        t = img.tensor
        out = CudaImage(_torch().zeros((int(height), int(width)) + tuple(t.shape[2:]), dtype=t.dtype, device=t.device))
    L.check(L.lib().cgs_image_warp(ctypes.byref(img._img), ctypes.byref(out._img), L.floats(np.asarray(H).reshape(-1)[:8]),
                                   _stream_ptr(stream)))
    return out


def fuse_images(ref, mov, est_H, true_H, channel=0, local=False, out=None, stream=None):
    """fuseAlignedImages (src/gpu/cudaImageFunctions.cuh:64-114) or, local=True,
    fuseAlignedImages_local (src/cpu/multispectral_ImageMatcher.cu:115-179):
    a 3-channel overlay of ref's size (zero-filled first, as the writers do)."""
    if out is None:
        out = CudaImage(_torch().zeros((ref.height(), ref.width(), 3), dtype=ref.tensor.dtype,
                                       device=ref.tensor.device))
    L.check(L.lib().cgs_image_fuse(ctypes.byref(ref._img), ctypes.byref(mov._img),
                                   L.floats(np.asarray(est_H).reshape(-1)[:8]),
                                   L.floats(np.asarray(true_H).reshape(-1)[:8]), int(channel), 1 if local else 0,
                                   ctypes.byref(out._img), _stream_ptr(stream)))
    return out


def write_png(path, img):
    """stbi_write_png of a uint8 CudaImage (or (H, W[, C]) uint8 array)."""
    a = img.tensor.cpu().numpy() if isinstance(img, CudaImage) else np.asarray(img)
    a = np.ascontiguousarray(a, np.uint8)
    c = 1 if a.ndim == 2 else a.shape[2]
    L.check(L.lib().cgs_write_png(os.fsencode(path), a.ctypes.data, a.shape[1], a.shape[0], c))


def read_png(path, req_comp=0):
    """stbi_load for a PNG file (cgs_read_png): a uint8 (H, W) array for one
    channel, (H, W, C) otherwise; req_comp 0 keeps the file's channels.
    Raises L.CgsError when the file cannot be read or decoded."""
    w, h, n = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    name = os.fsencode(path)
    L.check(L.lib().cgs_read_png(name, int(req_comp), None, 0, ctypes.byref(w), ctypes.byref(h), ctypes.byref(n)))
    c = int(req_comp) or n.value
    a = np.empty((h.value, w.value, c), np.uint8)
    L.check(L.lib().cgs_read_png(name, int(req_comp), a.ctypes.data, a.nbytes, None, None, None))
    return a[:, :, 0].copy() if c == 1 else a


def quantize_image(img, mod, stream=None):
    """transformPixel (src/cpu/multispectral_ImageMatcher.cu:411-430), in place
    on a uint8 CudaImage."""
    L.check(L.lib().cgs_image_quantize(ctypes.byref(img._img), float(mod), _stream_ptr(stream)))
    return img


def make_opts(functor, tie_mode=L.TIE_LOWEST):
    o = L.SearchOpts()
    L.lib().cgs_default_opts(ctypes.byref(o))
    o.cost, o.sampler, o.param_mode, o.mi_bins = functor.cost, functor.sampler, functor.param_mode, functor.bins
    o.tie_mode = tie_mode
    return o


class CudaGridSearcher:
    """CudaGridSearcher<float, float, D> (src/gpu/cudaGridSearch.cuh:540-855)."""

    def __init__(self, grid, result):
        self._grid = grid
        self._result = result

    def search(self, functor, img_moved, img_fixed, **kw):
        return self.search_by_value(functor, img_moved, img_fixed, **kw)

    def search_by_value(self, functor, img_moved, img_fixed, tie_mode=L.TIE_LOWEST, stream=None):
        """Fills the result tensor with the functor's cost at every grid point
        and returns the fused argmin (value, index)."""
        if self._result.size() != self._grid.numElements():
            raise ValueError("result tensor size differs from the grid's point count")
        o = make_opts(functor, tie_mode)
        from . import tuner
        o.tile_gpt = tuner.tuned_gpt(functor, img_moved, img_fixed)
        out = L.Argmin()
        L.check(L.lib().cgs_search(ctypes.byref(self._grid._g), ctypes.byref(img_moved._img),
                                   ctypes.byref(img_fixed._img), ctypes.byref(o),
                                   ctypes.c_void_p(self._result.data.data_ptr()), ctypes.byref(out),
                                   _stream_ptr(stream)))
        self._result.last_argmin = (out.value, out.index)
        return out.value, out.index

    def covariance_by_value(self, covar=None, stream=None):
        """CudaGridSearcher::covariance_by_value (src/gpu/cudaGridSearch.cuh:768-794)
        over the result tensor of the last search: adds the dim x dim
        covariance of the cost surface around its minimum into covar (a
        float32 numpy array, created as zeros when None) and returns it."""
        d = self._grid.getDimension()
        c = np.zeros((d, d), np.float32)
        m = L.Argmin()
        L.check(L.lib().cgs_covariance(ctypes.byref(self._grid._g), ctypes.c_void_p(self._result.data.data_ptr()),
                                       c.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), ctypes.byref(m),
                                       _stream_ptr(stream)))
        if covar is None:
            return c
        covar += c
        return covar

    # The reference's launch-policy variants compute the same costs; here they
    # all run the fused kernel.  STREAM_BLOCK_DIM / numCores / numThreads are
    # accepted for signature compatibility
    # (src/gpu/cudaGridSearch.cuh:562-649, 796-827).
    def search_by_value_stream(self, functor, STREAM_BLOCK_DIM, numCores, img_moved, img_fixed, **kw):
        return self.search_by_value(functor, img_moved, img_fixed, **kw)

    def search_by_value_block(self, functor, numThreads, img_moved, img_fixed, **kw):
        return self.search_by_value(functor, img_moved, img_fixed, **kw)

    def search_by_reference(self, functor, img_moved, img_fixed, **kw):
        return self.search_by_value(functor, img_moved, img_fixed, **kw)

    def search_by_reference_stream(self, functor, STREAM_BLOCK_DIM, numCores, img_moved, img_fixed, **kw):
        return self.search_by_value(functor, img_moved, img_fixed, **kw)


def eval_point(functor, img_moved, img_fixed, params, stream=None):
    """One functor evaluation at params; returns (cost, histogram or None)."""
    o = make_opts(functor)
    cost = ctypes.c_float()
    hist = None
    hp = None
    if functor.cost in L.MI_COSTS:
        hist = np.zeros(functor.bins * functor.bins, np.uint32)
        hp = hist.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
    p = [float(v) for v in params]
    L.check(L.lib().cgs_eval_point(ctypes.byref(img_moved._img), ctypes.byref(img_fixed._img), ctypes.byref(o),
                                   len(p), L.floats(p), ctypes.byref(cost), hp, _stream_ptr(stream)))
    return cost.value, (hist.reshape(functor.bins, functor.bins) if hist is not None else None)


class BatchSearcher:
    """Batched ragged search (one persistent kernel over (pair, grid-tile)
    work items, optionally replayed from a hipGraph) -- no reference
    counterpart; each pair's results equal CudaGridSearcher.search_by_value.

        b = BatchSearcher()
        b.plan([(functor, grid, img_mov, img_fix, result_tensor_or_None), ...], tiles=None)
        b.run(use_graph=True, stream=s)
        argmins = b.results(stream=s)          # [(value, index)] per pair
    """

    def __init__(self):
        h = ctypes.c_void_p()
        L.check(L.lib().cgs_batch_create(ctypes.byref(h)))
        self._h = h
        self._n = 0
        self._keep = None

    @staticmethod
    def _pair_array(pairs, tie_mode):
        arr = (L.BatchPair * len(pairs))()
        keep = []
        for k, (fn, grid, mov, fix, res) in enumerate(pairs):
            arr[k].moved = mov._img
            arr[k].fixed = fix._img
            arr[k].grid = grid._g
            arr[k].opts = make_opts(fn, tie_mode)
            arr[k].result = res.data.data_ptr() if res is not None else None
            keep.append((mov, fix, res))
        return arr, keep

    def plan(self, pairs, tiles=None, tie_mode=L.TIE_LOWEST):
        arr, keep = self._pair_array(pairs, tie_mode)
        t = None
        if tiles is not None:
            t = L.TileConfig(*[int(tiles.get(k, 0)) for k in ("fast_gpt", "mi_gpt", "wg_per_cu", "items_per_claim")])
        L.check(L.lib().cgs_batch_plan(self._h, arr, len(pairs), ctypes.byref(t) if t is not None else None))
        self._n = len(pairs)
        self._keep = (arr, keep)

    def run(self, use_graph=False, stream=None):
        L.check(L.lib().cgs_batch_run(self._h, 1 if use_graph else 0, _stream_ptr(stream)))

    def results(self, stream=None):
        out = (L.Argmin * self._n)()
        L.check(L.lib().cgs_batch_results(self._h, out, _stream_ptr(stream)))
        return [(a.value, a.index) for a in out]

    def refine(self, pairs, centres, radius=1, max_steps=8, threshold=float("inf"), fallback=True,
               tie_mode=L.TIE_LOWEST, stream=None):
        """Local refinement around per-pair centres with full-grid fallback
        (cgs_batch_refine).  pairs: (functor, grid, moved, fixed, ignored);
        centres: n x dim sample indices, or n x starts x dim for several
        descents per pair.  Returns a numpy record array with fields index,
        value, steps, outcome, local_index, local_value, points.  Drops any
        plan the batch held."""
        arr, keep = self._pair_array(pairs, tie_mode)
        c = np.ascontiguousarray(centres, np.int32)
        dim = pairs[0][1].getDimension()
        if c.ndim == 2:
            c = c[:, None, :]
        if c.ndim != 3 or c.shape[0] != len(pairs) or c.shape[2] != dim:
            raise ValueError("centres must be n x dim or n x starts x dim")
        c = np.ascontiguousarray(c)
        o = L.RefineOpts(int(radius), int(max_steps), float(threshold), 1 if fallback else 0, int(c.shape[1]))
        out = (L.RefineResult * len(pairs))()
        L.check(L.lib().cgs_batch_refine(self._h, arr, len(pairs), c.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                         ctypes.byref(o), out, _stream_ptr(stream)))
        self._n = 0
        self._keep = None
        del keep
        return np.array([(r.index, r.value, r.steps, r.outcome, r.local_index, r.local_value, r.points) for r in out],
                        dtype=[("index", np.int32), ("value", np.float32), ("steps", np.int32), ("outcome", np.int32),
                               ("local_index", np.int32), ("local_value", np.float32), ("points", np.int64)])

    def info(self):
        a, b, c = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        L.check(L.lib().cgs_batch_info(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return {"items": a.value, "points": b.value, "pixel_evals": c.value}

    def __del__(self):
        try:
            if self._h:
                L.lib().cgs_batch_destroy(self._h)
        except Exception:
            pass
