"""The matrix-product restatement (CudaMatrix::multiply's value: the fma
chain over k ascending from 0) pinned by known answers: the identity, exact
integer products (against numpy), agreement with binary64 numpy on random
data, and a cancellation case whose value fixes the summation order."""
import numpy as np

from oracle import oracle as O


def test_identity_and_integer_products():
    rng = np.random.default_rng(4)
    A = rng.integers(-50, 50, (7, 5)).astype(np.float32)
    assert np.array_equal(O.matrix_multiply(A, np.eye(5, dtype=np.float32)), A)
    B = rng.integers(-50, 50, (5, 9)).astype(np.float32)
    assert np.array_equal(O.matrix_multiply(A, B), (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float32))


def test_random_products_close_to_binary64():
    rng = np.random.default_rng(5)
    for dt, tol in ((np.float32, 1e-5), (np.float64, 1e-13)):
        A, B = rng.standard_normal((13, 37)).astype(dt), rng.standard_normal((37, 11)).astype(dt)
        want = A.astype(np.float64) @ B.astype(np.float64)
        got = O.matrix_multiply(A, B)
        assert got.dtype == dt and np.allclose(got, want, rtol=tol, atol=tol * 40)


def test_summation_order_is_k_ascending_fma():
    # 1e8 + 1 loses the 1 in binary32 (fma rounds after each step); then -1e8
    # cancels: 0.  Summing the 1 last, or pairwise, would give 1.
    A = np.array([[1e8, 1, -1e8]], np.float32)
    B = np.ones((3, 1), np.float32)
    assert O.matrix_multiply(A, B)[0, 0] == 0.0
    A2 = np.array([[1e8, -1e8, 1]], np.float32)
    assert O.matrix_multiply(A2, B)[0, 0] == 1.0
    # each step is one fma: (1 + 2^-12)^2 = 1 + 2^-11 + 2^-24 enters the sum
    # exactly, where a rounded product would lose the 2^-24
    a = np.float32(1 + 2 ** -12)
    A3 = np.array([[1, a]], np.float32)
    B3 = np.array([[np.float32(-1 - 2 ** -11)], [a]], np.float32)
    assert O.matrix_multiply(A3, B3)[0, 0] == np.float32(2.0 ** -24)
