"""GPU parity of the multispectral matcher: the image filter and quantisation
kernels bit-exact against the oracle, the Python pipeline run end to end on a
small PNG pair, its stdout parsed with the MATLAB harness's own patterns
(matlab/experiment_00.m:87-97) and its minimum checked against the oracle's
calcMIAlt search of the same grid, and the native CLI (cgs_matcher) run on
the same files as the Python CLI with identical output."""
import io
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from cugridsearch_amd import api as A
from cugridsearch_amd import _lib as L
from cugridsearch_amd import matcher as M
from cugridsearch_amd import synth

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("dtype,channels", [(np.uint8, 1), (np.uint8, 3), (np.float32, 1)])
def test_filter_matches_oracle(gpu, oracle, dtype, channels):
    rng = np.random.default_rng(3 + channels)
    shape = (37, 52) if channels == 1 else (37, 52, channels)
    img = (rng.random(shape) * 255).astype(dtype)
    filters = [M.UNSHARP_3X3, np.ones((1, 1), np.float32), np.full((5, 5), np.float32(1) / np.float32(25)),
               np.full((10, 10), np.float32(1) / np.float32(100)), rng.random((3, 4)).astype(np.float32) - 0.5]
    for filt in filters:
        got = A.filter_image(A.CudaImage.from_numpy(img), filt).tensor.cpu().numpy()
        want = oracle.image_filter(img, filt)
        assert got.dtype == want.dtype and np.array_equal(got.view(np.uint8), want.view(np.uint8)), filt.shape


def test_quantize_matches_oracle(gpu, oracle):
    v = np.tile(np.arange(256, dtype=np.uint8), 40).reshape(40, 256)
    for mod in (1.0, 2.5, 10.0, 0.3, 7.25, 100.0):
        img = A.CudaImage.from_numpy(v)
        A.quantize_image(img, mod)
        assert np.array_equal(img.tensor.cpu().numpy(), oracle.image_quantize(v, mod)), mod


def test_matcher_cli_protocol_and_minimum(gpu, oracle, tmp_path):
    from PIL import Image
    m, f, _ = synth.make_pair(21, 48, 40, sigma=2.0, perturb=2.0, dtype=np.uint8, gamma=True, noise=3.0)
    Image.fromarray(f).save(tmp_path / "fixed_0007.png")
    Image.fromarray(np.stack([m, m // 2, 255 - m], -1)).save(tmp_path / "moving.png")   # colour: stb's luma
    out = io.StringIO()
    res = M.run(M.load_gray(str(tmp_path / "fixed_0007.png")), M.load_gray(str(tmp_path / "moving.png")), out=out,
                results_file=str(tmp_path / "imageMatcherResults.txt"),
                names=(str(tmp_path / "fixed_0007.png"), str(tmp_path / "moving.png")))
    text = out.getvalue()
    # the harness: the first four {...} are the grid's start, end, step and count rows
    groups = re.findall(r"\{(.*?)\}", text, re.S)
    rows = [np.array([float(t) for t in gr.split(",")], np.float32) for gr in groups[:4]]
    num, start, end = M.coarse_grid(48, 40, 48, 40)
    assert np.allclose(rows[0], start, rtol=1e-5) and np.allclose(rows[1], end, rtol=1e-5)
    assert np.allclose(rows[3], num, rtol=1e-5)
    pat = r"Answer\[([^\]]*)]\s+Scale_factor\[([^\]]*)]\s+Homography_to_parameters\[([^\]]*)]\s+covar_matrix\[([^\]]*)]"
    mt = re.search(pat, text)
    assert mt, text[-800:]
    answer = np.array([float(t) for t in mt.group(1).split(",")])
    assert answer.shape == (8,) and np.allclose(answer, res["homography"], rtol=1e-5, atol=1e-6)
    assert [float(t) for t in mt.group(2).split(",")] == [1.0, 1.0]
    assert len(mt.group(3).split(",")) == 8 and len(mt.group(4).split(",")) == 64
    # the minimum: the oracle's calcMIAlt over the same grid (moving image filtered, fixed image not)
    mov = res["moving_filtered"].tensor.cpu().numpy()
    fix = M.load_gray(str(tmp_path / "fixed_0007.png"))
    assert np.array_equal(mov, oracle.image_filter(M.load_gray(str(tmp_path / "moving.png")), M.UNSHARP_3X3))
    c2, r = oracle.grid_update(start, end, num)
    want = oracle.search(mov, fix, oracle.COST_MI_ALT, oracle.BILINEAR, np.float32(start), r, c2,
                         mode=oracle.PARAM_NORM, bins=64)
    vo, io_ = oracle.argmin(want, oracle.TIE_REFERENCE)
    v, i = res["argmin"]
    assert abs(want[i] - vo) <= 2e-6 * abs(vo), (v, i, vo, io_)
    assert "Minimum found at point p = { " in text and "Covariance matrix:" in text
    line = (tmp_path / "imageMatcherResults.txt").read_text()
    assert line.startswith("input files,") and "Homography," in line and line.endswith("\n")


# ------------------------------------------------- the native CLI (cgs_matcher)

PROTOCOL = r"Answer\[([^\]]*)]\s+Scale_factor\[([^\]]*)]\s+Homography_to_parameters\[([^\]]*)]\s+covar_matrix\[([^\]]*)]"


def _parse_protocol(text):
    """What matlab/experiment_00.m:87-97 reads: the first four brace groups
    (the grid's start, end, step and count rows) and the four bracketed
    lists."""
    groups = re.findall(r"\{(.*?)\}", text, re.S)
    rows = [np.array([float(t) for t in gr.split(",")], np.float32) for gr in groups[:4]]
    mt = re.search(PROTOCOL, text)
    assert mt, text[-800:]
    lists = [np.array([float(t) for t in mt.group(k).split(",")], np.float32) for k in range(1, 5)]
    return rows, lists


def _run_cli(cmd, cwd):
    env = dict(os.environ)
    env["PYTHONPATH"] = ROOT + os.pathsep + env.get("PYTHONPATH", "")
    print("running:", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=cwd, env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-2000:])
    return r


@pytest.mark.parametrize("case", ["plain", "resize_gt_mod"])
def test_native_cli_matches_python_cli(gpu, tmp_path, case):
    """cgs_matcher (C++ over the C-ABI) and the Python CLI on the same files:
    the same stdout line for line (the wall-clock line aside), the same
    results line, the same output and resized images; and the C++ output
    parses with the MATLAB patterns into the grid and answer expected."""
    from PIL import Image
    from cugridsearch_amd import build as B
    assert os.path.exists(B.CLI), "cgs_matcher is not built (run build())"
    m, f, _ = synth.make_pair(23, 48, 40, sigma=2.0, perturb=2.0, dtype=np.uint8, gamma=True, noise=3.0)
    data = tmp_path / "data"
    data.mkdir()
    extra = []
    if case == "plain":
        Image.fromarray(f).save(data / "fixed_0007.png")
        Image.fromarray(np.stack([m, m // 2, 255 - m], -1)).save(data / "moving.png")
    else:
        # a fixed image of four times the area: resized to the moving image's size
        big = np.kron(f, np.ones((2, 2), np.uint8))
        Image.fromarray(big).save(data / "fixed_0007.png")
        Image.fromarray(m).save(data / "moving.png")
        (data / "gt.csv").write_text("id,h11\n3,1,1,1,1,1,1,1,1\n7,1.25,0.5,-2\n")   # a short row: wraps
        extra = ["-m", "2", "--gt", str(data / "gt.csv")]
    args = ["--i_ref", str(data / "fixed_0007.png"), "--i_mov", str(data / "moving.png"), "-o", "out.png",
            "-f", "fused.png", "--results-file", "results.txt", *extra]
    runs = {}
    for name, cmd in (("cpp", [B.CLI]), ("py", [sys.executable, "-m", "cugridsearch_amd.matcher"])):
        wd = tmp_path / name
        wd.mkdir()
        runs[name] = (_run_cli(cmd + args, str(wd)), wd)
    (rc, wc), (rp, wp) = runs["cpp"], runs["py"]

    def lines(text):
        # the wall clock differs run to run, and the device name comes from the
        # HIP runtime each process loads (torch's own for the Python CLI)
        out = [ln for ln in text.splitlines() if not ln.startswith("Total Time Taken: ")]
        assert re.fullmatch(r'GPU Device 0: ".*" with compute capability \d+\.\d+', out[0]), out[0]
        return [re.sub(r'".*"', '"<device>"', out[0])] + out[1:]
    lc, lp = lines(rc.stdout), lines(rp.stdout)
    for k, (a, b) in enumerate(zip(lc, lp)):
        assert a == b, (k, a, b)
    assert len(lc) == len(lp)
    assert sum(ln.startswith("Total Time Taken: ") for ln in rc.stdout.splitlines()) == 1
    assert (wc / "results.txt").read_text() == (wp / "results.txt").read_text()
    for img in ("out.png", "fused.png") + (("image_fixed_resized.png",) if case != "plain" else ()):
        a, b = A.read_png(str(wc / img)), A.read_png(str(wp / img))
        assert a.shape == b.shape and np.array_equal(a, b), img
    assert "Output image filename is out.png." in rc.stderr
    # the MATLAB parse of the native output
    rows, (answer, scale, params, covar) = _parse_protocol(rc.stdout)
    sf = (np.float32(1), np.float32(1)) if case == "plain" else (np.float32(0.5), np.float32(0.5))
    num, start, end = M.coarse_grid(48, 40, 48, 40, *sf)
    assert np.allclose(rows[0], start, rtol=1e-5) and np.allclose(rows[1], end, rtol=1e-5)
    assert np.allclose(rows[3], num, rtol=1e-5)
    assert answer.shape == (8,) and params.shape == (8,) and covar.shape == (64,)
    assert np.allclose(scale, sf)
    # the answer is the homography of the printed minimum (both printed with
    # six significant digits, hence the tolerance)
    pm = re.search(r"Minimum found at point p = \{ (.*?) \}", rc.stdout).group(1)
    best = [np.float32(float(t)) for t in pm.split(",")]
    want = M.parameters_to_homography_norm(best, 48, 40, 48, 40)
    assert np.allclose(answer, np.array(want, np.float32), rtol=2e-4, atol=1e-5)
    line = (wc / "results.txt").read_text()
    if case != "plain":
        assert "Rescaling" in rc.stderr
        gt = [np.float32(v) for v in (1.25, 0.5, -2, 7, 1.25, 0.5, -2, 7)]
        gt = [np.float32(v * np.float32(0.5)) for v in gt[:6]] + gt[6:]
        assert "GT Homography," + "".join(M.g(v) + "," for v in gt) in line
