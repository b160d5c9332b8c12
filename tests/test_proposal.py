"""CPU checks of the learned-proposal pieces: the network oracle pinned by
finite differences and closed forms, the homography geometry pinned by the
oracle's restatement of the reference's parametersToHomography, grid
sub-boxes through the C-ABI's host functions, and threshold calibration."""
import numpy as np
import pytest

from cugridsearch_amd import api as A
from cugridsearch_amd import proposal as P
from cugridsearch_amd import synth


def he_params(seed, fc_std=None):
    rng = np.random.default_rng(seed)
    from oracle import oracle as O
    p = np.zeros(O.cnn_param_count())
    cin, cout, o = [2, 32, 32, 64, 64, 128], [32, 32, 64, 64, 128, 128], 0
    for l in range(6):
        nw = cout[l] * 9 * cin[l]
        p[o:o + nw] = rng.normal(0, np.sqrt(2 / (9 * cin[l])), nw)
        p[o + nw:o + nw + cout[l]] = rng.normal(0, 0.05, cout[l])
        o += nw + cout[l]
    p[o:o + 1024] = rng.normal(0, fc_std or np.sqrt(1 / 128), 1024)
    p[o + 1024:] = rng.normal(0, 0.1, 8)
    return p


def test_cnn_oracle_gradient_matches_finite_differences(oracle):
    """The exact (binary64, no bf16 rounding) gradient against central
    differences of the loss, for parameters of every tensor."""
    rng = np.random.default_rng(1)
    p = he_params(2)
    x = rng.normal(size=(2, 16, 16, 2))
    t = rng.normal(0, 0.5, (2, 8))
    _, _, g = oracle.cnn_loss_grad(p, x, t, emulate=False)
    bounds = [0]
    cin, cout = [2, 32, 32, 64, 64, 128], [32, 32, 64, 64, 128, 128]
    for l in range(6):
        bounds += [bounds[-1] + cout[l] * 9 * cin[l], bounds[-1] + cout[l] * 9 * cin[l] + cout[l]]
    bounds += [bounds[-1] + 1024, bounds[-1] + 1032]
    checked = 0
    for a, b in zip(bounds[:-1], bounds[1:]):
        for i in rng.choice(np.arange(a, b), 3, replace=False):
            h = 1e-6
            pp, pm = p.copy(), p.copy()
            pp[i] += h
            pm[i] -= h
            fd = (oracle.cnn_loss_grad(pp, x, t, emulate=False, want_grad=False)[0] -
                  oracle.cnn_loss_grad(pm, x, t, emulate=False, want_grad=False)[0]) / (2 * h)
            assert abs(g[i] - fd) <= 1e-6 * max(abs(fd), 1e-3), (i, g[i], fd)
            checked += 1
    assert checked == 3 * 14


def test_cnn_oracle_closed_forms(oracle):
    P_ = oracle.cnn_param_count()
    rng = np.random.default_rng(3)
    x = rng.normal(size=(3, 16, 16, 2))
    # zero weights: the output is the FC bias; the loss is its mean square error
    p = np.zeros(P_)
    p[-8:] = np.arange(8) * 0.25
    t = rng.normal(size=(3, 8))
    loss, out, g = oracle.cnn_loss_grad(p, x, t, emulate=False)
    assert np.array_equal(out, np.tile(p[-8:], (3, 1)))
    assert np.isclose(loss, ((out - t) ** 2).sum() / 24)
    assert np.allclose(g[-8:], 2 * (out - t).sum(0) / 24)
    # target equal to the output: zero loss, zero gradient
    p = he_params(4)
    _, out, _ = oracle.cnn_loss_grad(p, x, None, emulate=False, want_grad=False)
    loss, _, g = oracle.cnn_loss_grad(p, x, out, emulate=False)
    assert loss == 0.0 and not g.any()
    # positive homogeneity of ReLU nets: scaling the input by 2 with zero
    # conv biases scales (out - fc bias) by 2
    p = he_params(5)
    o = 0
    for cin, cout in zip([2, 32, 32, 64, 64, 128], [32, 32, 64, 64, 128, 128]):
        o += cout * 9 * cin
        p[o:o + cout] = 0
        o += cout
    _, o1, _ = oracle.cnn_loss_grad(p, x, None, emulate=False, want_grad=False)
    _, o2, _ = oracle.cnn_loss_grad(p, 2 * x, None, emulate=False, want_grad=False)
    assert np.allclose(o2 - p[-8:], 2 * (o1 - p[-8:]), rtol=1e-12, atol=1e-12)


def test_round_bf16_known_values(oracle):
    one = 1.0
    cases = {one: one, 1 + 2 ** -8: one, 1 + 3 * 2 ** -9: 1 + 2 ** -7, 1 + 2 ** -7: 1 + 2 ** -7,
             -2.5: -2.5, 0.0: 0.0, 1 + 2 ** -9: one, 3.0e38: 2.0 ** 127 * (1 + 98 / 128)}
    for v, want in cases.items():
        got = oracle.round_bf16(np.array([v]))[0]
        assert np.float32(got) == np.float32(want), (v, got, want)
        assert oracle.cnn_lib().cnn_round_bf16(v) == got


def test_emulated_oracle_stays_near_exact(oracle):
    rng = np.random.default_rng(6)
    p = he_params(7)
    x = oracle.round_bf16(rng.normal(size=(2, 32, 32, 2)))
    t = rng.normal(0, 0.5, (2, 8))
    _, e_out, e_g = oracle.cnn_loss_grad(p, x, t, emulate=True)
    _, x_out, x_g = oracle.cnn_loss_grad(p, x, t, emulate=False)
    assert np.linalg.norm(e_out - x_out) <= 0.03 * np.linalg.norm(x_out)
    assert np.linalg.norm(e_g - x_g) <= 0.1 * np.linalg.norm(x_g)
    assert not np.array_equal(e_out, x_out)   # the rounding is applied


def test_homography_matches_reference_expression(oracle):
    """FP64 matrix product against the oracle's FP32 restatement of the
    reference's expanded expressions; decomposition and corner round trips."""
    rng = np.random.default_rng(8)
    for _ in range(100):
        p = (P.GRID_START + rng.random(8) * (P.GRID_END - P.GRID_START)).astype(np.float32)
        H = P.homography(p, 128.0, 128.0)
        Ho = np.append(np.asarray(oracle.homography(p, 128.0, 128.0), np.float64), 1.0).reshape(3, 3)
        assert np.abs(H - Ho).max() <= 1e-4 * np.abs(H).max()
        assert np.allclose(P.decompose(H, 128.0, 128.0), p, rtol=1e-9, atol=1e-12)
        assert np.allclose(P.offsets_to_params(P.corner_offsets(H, 256, 256), 256, 256), p, rtol=1e-7, atol=1e-10)
    assert np.allclose(P.homography([0, 1, 1, 0, 0, 0, 0, 0], 128.0, 128.0), np.eye(3))
    H = P.homography([0, 1, 1, 0, 3, -2, 0, 0], 128.0, 128.0)
    assert np.allclose(P.corner_offsets(H, 256, 256), [3, -2] * 4)


def test_generator_corners_match_dlt():
    """synth's H_true maps the corners to the perturbed corners, so its
    corner offsets are the perturbations (within +-32 px)."""
    m, f, H = synth.make_pair(P.TRAIN_SEED + 3, 256, sigma=P.SIGMA, perturb=P.PERTURB, dtype=np.float32,
                              extra_shapes=True)
    off = P.corner_offsets(H, 256, 256)
    assert np.abs(off).max() <= 32.0 and m.shape == f.shape == (256, 256)


def test_subbox_points_are_grid_points():
    g = P.grid()
    box = g.subbox([1, 2, 0, 3, 4, 5, 2, 0], [3, 2, 6, 1, 2, 1, 3, 2])
    assert box.numElements() == 3 * 2 * 6 * 1 * 2 * 1 * 3 * 2
    assert list(box.offset()) == [1, 2, 0, 3, 4, 5, 2, 0]
    count = g.getAxisSampleCounts().astype(int)
    inc = np.cumprod(np.r_[1, count[:-1]])
    bc = box.getAxisSampleCounts().astype(int)
    for k in range(box.numElements()):
        full = int(((P.index_axes(k, bc) + box.offset()) * inc).sum())
        assert np.array_equal(box.getGridPoint(k).view(np.uint32), g.getGridPoint(full).view(np.uint32))
    inner = box.subbox([1, 0, 2, 0, 1, 0, 0, 1], [1, 1, 2, 1, 1, 1, 2, 1])
    assert list(inner.offset()) == [2, 2, 2, 3, 5, 5, 2, 1]


@pytest.mark.parametrize("lo,cnt", [([0] * 7 + [-1], [1] * 8), ([5] * 8, [2] * 8), ([0] * 8, [0] + [1] * 7)])
def test_subbox_rejects_boxes_outside(lo, cnt):
    from cugridsearch_amd import _lib as L
    with pytest.raises(L.CgsError):
        P.grid().subbox(lo, cnt)


def test_nearest_centre_and_index_axes():
    g = P.grid()
    start, res, count = g.start_point(), g.resolution(), g.getAxisSampleCounts()
    for i in (0, 12345, g.numElements() - 1):
        ax = P.index_axes(i, count)
        assert np.array_equal(P.nearest_centre(g.getGridPoint(i), start, res, count), ax)
    far = P.nearest_centre(np.full(8, 1e3), start, res, count)
    assert np.array_equal(far, count.astype(int) - 1)


def test_calibrate_threshold():
    v = np.array([-0.99, -0.98, -0.97, -0.96, -0.95, -0.9, -0.94, -0.93, -0.92, -0.91])
    assert P.calibrate_threshold(v, 0.0) == np.inf
    assert P.calibrate_threshold(v, 0.1) == -0.91          # the worst tenth falls back
    assert P.calibrate_threshold(v, 0.3) == -0.93
    assert (v > P.calibrate_threshold(v, 0.2)).sum() == 2
    assert P.calibrate_threshold(v, 1.0) == -0.99
    assert P.calibrate_threshold([], 0.1) == np.inf


def test_network_layout():
    from cugridsearch_amd import cnn
    lay = cnn.layout()
    assert lay[0] == ("conv1.weight", 0, (32, 3, 3, 2))
    assert lay[-1][0] == "fc.bias" and lay[-1][1] + 8 == 287752
    p = cnn.init_params(0)
    assert p.dtype == np.float32 and p.size == 287752
    assert np.allclose(np.std(p[:576]), np.sqrt(2 / 18), rtol=0.15)
