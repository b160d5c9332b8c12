"""GPU parity of the tensor utilities (cgs_tensor_fill, cgs_tensor_binary,
cgs_matrix_multiply) and of the C++ facade's CudaMatrix / CudaTensor
operations: fill and element-wise results equal numpy's element for element
(uint8 wrapping modulo 256), products equal the oracle's fma chain bit for
bit."""
import ctypes
import json
import os
import subprocess

import numpy as np
import pytest
import torch

from cugridsearch_amd import _lib as L
from cugridsearch_amd.build import DRIVER

pytestmark = pytest.mark.gpu

DT = {np.float32: L.DT_F32, np.float64: L.DT_F64, np.int32: L.DT_I32, np.uint8: L.DT_U8}


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("dt", [np.float32, np.float64, np.int32, np.uint8])
def test_fill_and_elementwise(gpu, dt):
    rng = np.random.default_rng(3)
    n = 100_003
    t = torch.zeros(n, dtype=_dev(np.zeros(1, dt)).dtype, device="cuda")
    L.check(L.lib().cgs_tensor_fill(ctypes.c_void_p(t.data_ptr()), n, DT[dt], 7.0, None))
    assert (t.cpu().numpy() == dt(7)).all()
    hi = 256 if dt == np.uint8 else 1000
    for op, fn in ((L.OP_ADD, np.add), (L.OP_SUB, np.subtract), (L.OP_MUL, np.multiply)):
        a = rng.integers(0, hi, n).astype(dt)
        b = rng.integers(0, hi, n).astype(dt)
        if dt in (np.float32, np.float64):
            a, b = a + dt(0.37), b * dt(1.5)
        da, db = _dev(a), _dev(b)
        L.check(L.lib().cgs_tensor_binary(op, ctypes.c_void_p(da.data_ptr()), ctypes.c_void_p(db.data_ptr()), n, DT[dt],
                                          None))
        with np.errstate(over="ignore"):
            want = fn(a, b).astype(dt)
        assert np.array_equal(da.cpu().numpy(), want), (dt, op)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_matrix_multiply_matches_oracle(gpu, oracle, dt):
    rng = np.random.default_rng(8)
    for m, k, n in ((1, 1, 1), (3, 4, 2), (17, 33, 9), (64, 130, 70), (5, 1000, 3)):
        A, B = rng.standard_normal((m, k)).astype(dt), rng.standard_normal((k, n)).astype(dt)
        dA, dB = _dev(A), _dev(B)
        dC = torch.empty((m, n), dtype=dA.dtype, device="cuda")
        L.check(L.lib().cgs_matrix_multiply(ctypes.c_void_p(dA.data_ptr()), ctypes.c_void_p(dB.data_ptr()),
                                            ctypes.c_void_p(dC.data_ptr()), m, k, n, DT[dt], None))
        got = dC.cpu().numpy()
        assert np.array_equal(got.view(np.uint8), oracle.matrix_multiply(A, B).view(np.uint8)), (m, k, n)


def test_facade_matrix_and_tensor_ops(gpu, oracle):
    if not os.path.exists(DRIVER):
        pytest.fail(f"{DRIVER} missing: run __graft_entry__.build()")
    r = subprocess.run([DRIVER, "--matrix"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Matrix C 3 x 2 elements of f" in r.stdout and "{ " in r.stdout
    res = json.loads([l for l in r.stdout.splitlines() if l.startswith("{\"A\"")][-1])
    A = np.eye(3, 4, dtype=np.float32) * 2
    A[2] = [1, -2, 3, 0.5]
    A[:, 3] = [4, 5, 6]
    B = np.arange(1, 9, dtype=np.float32).reshape(4, 2)
    C = oracle.matrix_multiply(A, B)
    assert np.array_equal(np.array(res["A"], np.float32).reshape(3, 4), A)
    assert np.array_equal(np.array(res["C"], np.float32).reshape(3, 2), C)
    T = (((C + C) * C) - C).astype(np.float32)
    assert np.array_equal(np.array(res["T"], np.float32).reshape(3, 2), T)
    assert res["u"] == [144.0] * 5
