"""The proposal network over the C-ABI (cgs_cnn_* in include/cgs.h).

No reference counterpart.  Every pass runs in libcgs.so (hand-written
BF16 matrix-core kernels); this module holds the handle, the parameter
initialisation and file I/O.  Device buffers are torch tensors (allocation
only): inputs are int16 tensors carrying bf16 bit patterns.
"""
import ctypes

import numpy as np

from . import _lib as L

CIN = (2, 32, 32, 64, 64, 128)
COUT = (32, 32, 64, 64, 128, 128)
STRIDE = (1, 2, 1, 2, 1, 2)
SIDE = 128
FEAT = 128
NOUT = 8


def param_count():
    return int(L.lib().cgs_cnn_param_count())


def layout():
    """[(name, offset, shape)] of every tensor in the flat parameter vector."""
    out, p = [], 0
    for l in range(6):
        shape = (COUT[l], 3, 3, CIN[l])
        out.append((f"conv{l + 1}.weight", p, shape))
        p += int(np.prod(shape))
        out.append((f"conv{l + 1}.bias", p, (COUT[l],)))
        p += COUT[l]
    out.append(("fc.weight", p, (NOUT, FEAT)))
    p += NOUT * FEAT
    out.append(("fc.bias", p, (NOUT,)))
    return out


def init_params(seed):
    """He-normal conv weights (std sqrt(2 / (9 Cin))), FC weight normal with
    std 0.1 / sqrt(128), zero biases; float32."""
    rng = np.random.default_rng(seed)
    p = np.zeros(param_count(), np.float32)
    for name, off, shape in layout():
        n = int(np.prod(shape))
        if name.startswith("conv") and name.endswith("weight"):
            p[off:off + n] = rng.normal(0.0, np.sqrt(2.0 / (9 * shape[3])), n)
        elif name == "fc.weight":
            p[off:off + n] = rng.normal(0.0, 0.1 / np.sqrt(FEAT), n)
    return p


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _sp(stream):
    if stream is None:
        return None
    return ctypes.c_void_p(stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream))


class ProposalNet:
    """Handle on one network instance (parameters, optimiser state and the
    activations of the last forward pass) for batches up to max_batch."""

    def __init__(self, max_batch=256, params=None, seed=None):
        h = ctypes.c_void_p()
        L.check(L.lib().cgs_cnn_create(int(max_batch), ctypes.byref(h)))
        self._h = h
        self.max_batch = int(max_batch)
        if params is not None:
            self.set_params(params)
        elif seed is not None:
            self.set_params(init_params(seed))

    def set_params(self, params):
        p = np.ascontiguousarray(params, np.float32)
        if p.size != param_count():
            raise ValueError(f"expected {param_count()} parameters, got {p.size}")
        L.check(L.lib().cgs_cnn_set_params(self._h, _fp(p)))

    def get_params(self):
        p = np.zeros(param_count(), np.float32)
        L.check(L.lib().cgs_cnn_get_params(self._h, _fp(p)))
        return p

    def get_grads(self):
        g = np.zeros(param_count(), np.float32)
        L.check(L.lib().cgs_cnn_get_grads(self._h, _fp(g)))
        return g

    @staticmethod
    def prepare(moved, fixed, out=None, stream=None):
        """Network inputs of pairs (lists of CudaImage, 128x128 or 256x256):
        int16 tensor [n, 128, 128, 2] of bf16 bit patterns."""
        import torch
        n = len(moved)
        if len(fixed) != n or n == 0:
            raise ValueError("moved and fixed need the same non-zero length")
        if out is None:
            out = torch.empty((n, SIDE, SIDE, 2), dtype=torch.int16, device="cuda")
        mi = (L.Image * n)(*[m._img for m in moved])
        fi = (L.Image * n)(*[f._img for f in fixed])
        L.check(L.lib().cgs_cnn_prepare(mi, fi, n, ctypes.c_void_p(out.data_ptr()), _sp(stream)))
        return out

    def forward(self, x, out=None, stream=None):
        """Corner offsets [n, 8] (float32, device) of inputs x [n, 128, 128, 2]."""
        import torch
        n = int(x.shape[0])
        if tuple(x.shape[1:]) != (SIDE, SIDE, 2) or x.dtype != torch.int16 or not x.is_contiguous():
            raise ValueError("x must be a contiguous int16 tensor [n, 128, 128, 2]")
        if out is None:
            out = torch.empty((n, NOUT), dtype=torch.float32, device=x.device)
        L.check(L.lib().cgs_cnn_forward(self._h, ctypes.c_void_p(x.data_ptr()), n, ctypes.c_void_p(out.data_ptr()),
                                        _sp(stream)))
        self._last_x = x   # the backward pass reads it
        return out

    def backward(self, target, stream=None, want_loss=True):
        """Gradient of the mean squared error of the last forward pass; the
        loss (float) when want_loss."""
        import torch
        if target.dtype != torch.float32 or not target.is_contiguous():
            raise ValueError("target must be a contiguous float32 tensor [n, 8]")
        loss = ctypes.c_float()
        L.check(L.lib().cgs_cnn_backward(self._h, ctypes.c_void_p(target.data_ptr()),
                                         ctypes.byref(loss) if want_loss else None, _sp(stream)))
        return loss.value if want_loss else None

    def adam_step(self, lr, beta1=0.9, beta2=0.999, eps=1e-8, stream=None):
        L.check(L.lib().cgs_cnn_adam_step(self._h, float(lr), float(beta1), float(beta2), float(eps), _sp(stream)))

    def save(self, path, **meta):
        np.savez(path, params=self.get_params(), **meta)

    @classmethod
    def load(cls, path, max_batch=256):
        with np.load(path, allow_pickle=False) as z:
            return cls(max_batch=max_batch, params=z["params"])

    def __del__(self):
        try:
            if self._h:
                L.lib().cgs_cnn_destroy(self._h)
        except Exception:
            pass
