"""ctypes binding of libcgs.so (the C-ABI declared in include/cgs.h).

The shared library is the product: nothing here computes a search result.
If libcgs.so is missing the import of any compute entry point fails loudly.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcgs.so")

CGS_OK = 0
STATUS_NAMES = {0: "CGS_OK", 1: "CGS_ERR_INVALID_ARGUMENT", 2: "CGS_ERR_GRID_TOO_LARGE", 3: "CGS_ERR_UNSUPPORTED",
                4: "CGS_ERR_HIP", 5: "CGS_ERR_NO_DEVICE", 6: "CGS_ERR_TOO_LARGE"}

COST_SAD, COST_NCC, COST_MI, COST_SQD, COST_SAD_SUM, COST_MI_PLAIN = range(6)
COST_NCC_STREAM, COST_SQD_STREAM, COST_SAD_STREAM = 6, 7, 8
COST_SQD_ALT, COST_NCC_ALT, COST_NCC_STREAM_ALT, COST_MI_STREAM_ALT, COST_MI_ALT = 9, 10, 11, 12, 13
MI_COSTS = (COST_MI, COST_MI_PLAIN, COST_MI_STREAM_ALT, COST_MI_ALT)
SAMPLER_NEAREST, SAMPLER_BILINEAR = 0, 1
PARAM_CENTRED, PARAM_NONCENTRED, PARAM_RAW_H, PARAM_TRANSLATION, PARAM_NORM = range(5)
TIE_LOWEST, TIE_REFERENCE = 0, 1
PIX_U8, PIX_F32 = 0, 1
DT_F32, DT_F64, DT_I32, DT_U8 = 0, 1, 2, 3
OP_ADD, OP_SUB, OP_MUL = 0, 1, 2
MAX_DIM = 8
ABI_VERSION = 6          # CGS_ABI_VERSION of include/cgs.h these layouts follow
FLAG_TIME_KERNEL = 1


class CgsError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


class Grid(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int32), ("start", ctypes.c_float * 8), ("end", ctypes.c_float * 8),
                ("count", ctypes.c_float * 8), ("resolution", ctypes.c_float * 8),
                ("increments", ctypes.c_int32 * 8), ("num_elements", ctypes.c_int64), ("offset", ctypes.c_int32 * 8)]


class Image(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("width", ctypes.c_int32), ("height", ctypes.c_int32),
                ("channels", ctypes.c_int32), ("pixtype", ctypes.c_int32)]


class SearchOpts(ctypes.Structure):
    _fields_ = [("cost", ctypes.c_int32), ("sampler", ctypes.c_int32), ("param_mode", ctypes.c_int32),
                ("tie_mode", ctypes.c_int32), ("mi_bins", ctypes.c_int32), ("flags", ctypes.c_int32),
                ("tile_gpt", ctypes.c_int32)]


class Argmin(ctypes.Structure):
    _fields_ = [("value", ctypes.c_float), ("index", ctypes.c_int32)]


class BatchPair(ctypes.Structure):
    _fields_ = [("moved", Image), ("fixed", Image), ("grid", Grid), ("opts", SearchOpts),
                ("result", ctypes.c_void_p)]


class TileConfig(ctypes.Structure):
    _fields_ = [("fast_gpt", ctypes.c_int32), ("mi_gpt", ctypes.c_int32), ("wg_per_cu", ctypes.c_int32),
                ("items_per_claim", ctypes.c_int32)]


class RefineOpts(ctypes.Structure):
    _fields_ = [("radius", ctypes.c_int32), ("max_steps", ctypes.c_int32), ("threshold", ctypes.c_float),
                ("fallback", ctypes.c_int32), ("starts", ctypes.c_int32)]


class RefineResult(ctypes.Structure):
    _fields_ = [("index", ctypes.c_int32), ("value", ctypes.c_float), ("steps", ctypes.c_int32),
                ("outcome", ctypes.c_int32), ("local_index", ctypes.c_int32), ("local_value", ctypes.c_float),
                ("points", ctypes.c_int64)]


REFINE_LOCAL, REFINE_FALLBACK_BOUNDARY, REFINE_FALLBACK_THRESHOLD = range(3)


# name -> (restype, argtypes); every symbol include/cgs.h declares.
_P = ctypes.POINTER
SIGNATURES = {
    "cgs_version": (ctypes.c_char_p, []),
    "cgs_abi_version": (ctypes.c_int32, []),
    "cgs_default_opts": (None, [_P(SearchOpts)]),
    "cgs_grid_init": (ctypes.c_int, [_P(Grid), ctypes.c_int32, _P(ctypes.c_float), _P(ctypes.c_float),
                                     _P(ctypes.c_float)]),
    "cgs_grid_point": (ctypes.c_int, [_P(Grid), ctypes.c_int64, _P(ctypes.c_float)]),
    "cgs_grid_subbox": (ctypes.c_int, [_P(Grid), _P(ctypes.c_int32), _P(ctypes.c_int32), _P(Grid)]),
    "cgs_search": (ctypes.c_int, [_P(Grid), _P(Image), _P(Image), _P(SearchOpts), ctypes.c_void_p, _P(Argmin),
                                  ctypes.c_void_p]),
    "cgs_search_noalloc": (ctypes.c_int, [_P(Grid), _P(Image), _P(Image), _P(SearchOpts), ctypes.c_void_p,
                                          _P(Argmin), ctypes.c_void_p]),
    "cgs_fma_peak": (ctypes.c_int, [_P(ctypes.c_float), ctypes.c_void_p]),
    "cgs_check_rcp": (ctypes.c_int, [_P(ctypes.c_uint32), ctypes.c_void_p]),
    "cgs_valu_rates": (ctypes.c_int, [_P(ctypes.c_float), ctypes.c_int, ctypes.c_void_p]),
    "cgs_find_extrema": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, _P(Argmin),
                                        ctypes.c_void_p]),
    "cgs_covariance": (ctypes.c_int, [_P(Grid), ctypes.c_void_p, _P(ctypes.c_float), _P(Argmin), ctypes.c_void_p]),
    "cgs_image_filter": (ctypes.c_int, [_P(Image), _P(Image), _P(ctypes.c_float), ctypes.c_int32, ctypes.c_int32,
                                        ctypes.c_void_p]),
    "cgs_image_filter_channels": (ctypes.c_int, [_P(Image), _P(Image), _P(ctypes.c_float), ctypes.c_int32,
                                                 ctypes.c_int32, ctypes.c_uint32, ctypes.c_void_p]),
    "cgs_image_quantize": (ctypes.c_int, [_P(Image), ctypes.c_float, ctypes.c_void_p]),
    "cgs_image_resize": (ctypes.c_int, [_P(Image), _P(Image), ctypes.c_void_p]),
    "cgs_tensor_fill": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_double, ctypes.c_void_p]),
    "cgs_tensor_binary": (ctypes.c_int, [ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32,
                                         ctypes.c_void_p]),
    "cgs_matrix_multiply": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                           ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p]),
    "cgs_image_warp": (ctypes.c_int, [_P(Image), _P(Image), _P(ctypes.c_float), ctypes.c_void_p]),
    "cgs_image_fuse": (ctypes.c_int, [_P(Image), _P(Image), _P(ctypes.c_float), _P(ctypes.c_float), ctypes.c_int32,
                                      ctypes.c_int32, _P(Image), ctypes.c_void_p]),
    "cgs_write_png": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    "cgs_read_png": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_int64,
                                    _P(ctypes.c_int32), _P(ctypes.c_int32), _P(ctypes.c_int32)]),
    "cgs_eval_point": (ctypes.c_int, [_P(Image), _P(Image), _P(SearchOpts), ctypes.c_int32, _P(ctypes.c_float),
                                      _P(ctypes.c_float), _P(ctypes.c_uint32), ctypes.c_void_p]),
    "cgs_batch_create": (ctypes.c_int, [_P(ctypes.c_void_p)]),
    "cgs_batch_plan": (ctypes.c_int, [ctypes.c_void_p, _P(BatchPair), ctypes.c_int32, _P(TileConfig)]),
    "cgs_batch_run": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p]),
    "cgs_batch_results": (ctypes.c_int, [ctypes.c_void_p, _P(Argmin), ctypes.c_void_p]),
    "cgs_batch_info": (ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_int64), _P(ctypes.c_int64),
                                      _P(ctypes.c_int64)]),
    "cgs_batch_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "cgs_batch_refine": (ctypes.c_int, [ctypes.c_void_p, _P(BatchPair), ctypes.c_int32, _P(ctypes.c_int32),
                                        _P(RefineOpts), _P(RefineResult), ctypes.c_void_p]),
    "cgs_cnn_param_count": (ctypes.c_int64, []),
    "cgs_cnn_create": (ctypes.c_int, [ctypes.c_int32, _P(ctypes.c_void_p)]),
    "cgs_cnn_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "cgs_cnn_set_params": (ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_float)]),
    "cgs_cnn_get_params": (ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_float)]),
    "cgs_cnn_prepare": (ctypes.c_int, [_P(Image), _P(Image), ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p]),
    "cgs_cnn_forward": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
                                       ctypes.c_void_p]),
    "cgs_cnn_backward": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, _P(ctypes.c_float), ctypes.c_void_p]),
    "cgs_cnn_get_grads": (ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_float)]),
    "cgs_cnn_adam_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                         ctypes.c_float, ctypes.c_void_p]),
    "cgs_last_kernel_ms": (ctypes.c_float, []),
    "cgs_last_error": (ctypes.c_char_p, []),
}

_lib = None


def lib():
    """Loads libcgs.so once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build() (hipcc for gfx950) first")
        # torch ships its own libamdhip64; loading it first makes libcgs.so
        # bind to that same HIP runtime (two runtimes in one process do not
        # share devices, streams or allocations).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        if L.cgs_abi_version() != ABI_VERSION:
            raise ImportError(f"{LIB_PATH} has ABI version {L.cgs_abi_version()}, this binding expects "
                              f"{ABI_VERSION}: rebuild with __graft_entry__.build()")
        _lib = L
    return _lib


def check(status):
    if status != CGS_OK:
        raise CgsError(status, lib().cgs_last_error().decode())
    return status


def floats(vals, n=None):
    vals = list(vals)
    if n is not None:
        vals = vals + [0.0] * (n - len(vals))
    return (ctypes.c_float * len(vals))(*vals)
