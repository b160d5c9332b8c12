"""GPU proposal network (BF16 matrix-core kernels, FP32 accumulation) against
the CPU FP64 oracle (oracle/cnn_oracle.c), through the C-ABI.

The oracle runs in its bf16-emulating mode: it rounds the weights, every
stored activation and every inter-layer gradient to bf16 where the GPU does,
so what remains is FP32-vs-FP64 accumulation (and the rare bf16 rounding or
ReLU mask flip that it causes).  Bars, written in each test: network output
within 1% (relative L2), every parameter tensor's gradient within 3%
(relative L2) with cosine similarity >= 0.999."""
import numpy as np
import pytest

from cugridsearch_amd import cnn as C
from cugridsearch_amd.api import CudaImage

pytestmark = pytest.mark.gpu

OUT_RTOL = 1e-2
GRAD_RTOL = 3e-2
GRAD_COS = 0.999


def bf16_input(rng, n, side=128):
    """Random network input as (int16 device tensor of bf16 bits, float64 values)."""
    import torch
    from oracle import oracle as O
    v = O.round_bf16(rng.normal(size=(n, side, side, 2)))
    bits = (v.astype(np.float32).view(np.uint32) >> 16).astype(np.uint16).view(np.int16)
    return torch.from_numpy(np.ascontiguousarray(bits)).cuda(), v


def params_with_biases(seed):
    rng = np.random.default_rng(seed + 100)
    p = C.init_params(seed)
    for name, off, shape in C.layout():
        if name.endswith("bias"):
            p[off:off + shape[0]] = rng.normal(0.0, 0.05, shape[0])
    p[[o for n, o, _ in C.layout() if n == "fc.weight"][0]:][:1024] *= 10.0   # FC weights of order 1/sqrt(128)
    return p


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def test_param_count_matches_oracle(gpu, oracle):
    assert C.param_count() == oracle.cnn_param_count() == 287752


def test_forward_matches_oracle(gpu, oracle):
    rng = np.random.default_rng(11)
    p = params_with_biases(1)
    X, xv = bf16_input(rng, 3)
    net = C.ProposalNet(max_batch=8, params=p)
    out = net.forward(X).cpu().numpy()
    _, emu, _ = oracle.cnn_loss_grad(p, xv, None, emulate=True, want_grad=False)
    _, exact, _ = oracle.cnn_loss_grad(p, xv, None, emulate=False, want_grad=False)
    e = rel(out, emu)
    print(f"forward: rel L2 vs bf16-emulating oracle {e:.2e}, vs exact FP64 {rel(out, exact):.2e}")
    assert e <= OUT_RTOL
    assert np.abs(out - emu).max() <= 2 * OUT_RTOL * np.abs(emu).max()


def test_gradients_match_oracle(gpu, oracle):
    import torch
    rng = np.random.default_rng(12)
    p = params_with_biases(2)
    X, xv = bf16_input(rng, 3)
    T = rng.normal(0.0, 0.5, (3, 8)).astype(np.float32)
    net = C.ProposalNet(max_batch=8, params=p)
    net.forward(X)
    loss = net.backward(torch.from_numpy(T).cuda())
    g = net.get_grads()
    lo, _, go = oracle.cnn_loss_grad(p, xv, T, emulate=True)
    assert abs(loss - lo) <= 2e-2 * lo
    for name, off, shape in C.layout():
        n = int(np.prod(shape))
        a, b = g[off:off + n].astype(np.float64), go[off:off + n]
        e = rel(a, b)
        cos = float(a @ b / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-30))
        print(f"{name:14s} rel {e:.2e} cos {cos:.6f}")
        assert e <= GRAD_RTOL, name
        assert cos >= GRAD_COS, name


def test_prepare_matches_numpy(gpu, oracle):
    import torch
    rng = np.random.default_rng(13)
    m8 = rng.integers(0, 256, (256, 256), dtype=np.uint8)
    f32 = (rng.random((128, 128)) * 255).astype(np.float32)
    flat = np.full((128, 128), 7.0, np.float32)
    X = C.ProposalNet.prepare([CudaImage.from_numpy(m8), CudaImage.from_numpy(flat)],
                              [CudaImage.from_numpy(f32), CudaImage.from_numpy(f32)]).cpu().numpy()

    def expect(img):
        a = img.astype(np.float32)
        if a.shape[0] == 256:
            a = ((a[0::2, 0::2] + a[0::2, 1::2]) + (a[1::2, 0::2] + a[1::2, 1::2])) * np.float32(0.25)
        d = a.astype(np.float64)
        mean = d.mean()
        var = (d * d).mean() - mean * mean
        inv = 1.0 / np.sqrt(var) if var > 1e-12 else 0.0
        v = oracle.round_bf16(((d - mean) * inv).astype(np.float32))
        return (v.astype(np.float32).view(np.uint32) >> 16).astype(np.uint16)

    for i, (mv, fx) in enumerate([(m8, f32), (flat, f32)]):
        for ch, img in ((0, mv), (1, fx)):
            got = X[i, :, :, ch].view(np.uint16)
            want = expect(img)
            diff = np.abs(got.astype(np.int32) - want.astype(np.int32))
            assert diff.max() <= 1 and (diff == 0).mean() >= 0.999, (i, ch)
    assert not X[1, :, :, 0].any()   # a flat image maps to 0


def test_adam_step_matches_numpy(gpu):
    import torch
    rng = np.random.default_rng(14)
    p = params_with_biases(3)
    X, _ = bf16_input(rng, 2)
    T = torch.from_numpy(rng.normal(0.0, 0.5, (2, 8)).astype(np.float32)).cuda()
    net = C.ProposalNet(max_batch=4, params=p)
    m = np.zeros_like(p, np.float64)
    v = np.zeros_like(p, np.float64)
    w = p.astype(np.float64)
    lr, b1, b2, eps = 1e-3, 0.9, 0.999, 1e-8
    for t in range(1, 4):
        net.forward(X)
        net.backward(T, want_loss=False)
        g = net.get_grads().astype(np.float64)
        net.adam_step(lr, b1, b2, eps)
        m = b1 * m + (1 - b1) * g
        v = b2 * v + (1 - b2) * g * g
        w = w - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
        got = net.get_params()
        assert np.allclose(got, w, rtol=1e-5, atol=1e-7), t


def test_training_fits_a_fixed_batch(gpu):
    import torch
    rng = np.random.default_rng(15)
    X, _ = bf16_input(rng, 16)
    T = torch.from_numpy(rng.normal(0.0, 0.5, (16, 8)).astype(np.float32)).cuda()
    net = C.ProposalNet(max_batch=16, seed=4)
    net.forward(X)
    first = net.backward(T)
    for _ in range(150):
        net.forward(X)
        last = net.backward(T)
        net.adam_step(1e-3)
    print(f"loss {first:.4f} -> {last:.4f}")
    assert last < 0.2 * first


def test_batch_bounds_are_checked(gpu):
    import torch
    from cugridsearch_amd import _lib as L
    net = C.ProposalNet(max_batch=2, seed=0)
    X = torch.zeros((3, 128, 128, 2), dtype=torch.int16, device="cuda")
    with pytest.raises(L.CgsError):
        net.forward(X)
    fresh = C.ProposalNet(max_batch=2, seed=0)
    with pytest.raises(L.CgsError):
        fresh.backward(torch.zeros((2, 8), dtype=torch.float32, device="cuda"))
