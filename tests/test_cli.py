"""The native matcher CLI (cgs_matcher) on the CPU: its argument handling and
the exits the reference takes before any device work
(src/cpu/multispectral_ImageMatcher.cu:455-514): no reference image, no
moving image, an image that fails to load.  The full pipeline runs in the GPU
tests (tests/test_gpu_matcher.py)."""
import os
import subprocess

import pytest

from cugridsearch_amd import build as B


@pytest.fixture(scope="module")
def cli():
    if not os.path.exists(B.LIB):
        pytest.skip("libcgs.so is not built")
    if not os.path.exists(B.CLI) or os.path.getmtime(B.CLI) < os.path.getmtime(B.LIB):
        B.build_cli()
    return B.CLI


def run(cli, *args, cwd=None):
    env = dict(os.environ)
    env["HIP_VISIBLE_DEVICES"] = ""
    return subprocess.run([cli, *args], capture_output=True, text=True, timeout=60, cwd=cwd, env=env)


def test_missing_images_exit_like_the_reference(cli):
    r = run(cli)
    assert r.returncode == 1
    assert "No input reference image filename was provided. Exiting.." in r.stderr
    assert "--i_ref arg" in r.stdout and "--fusedoutput arg" in r.stdout
    r = run(cli, "--i_ref", "a.png")
    assert r.returncode == 1 and "No input moving image filename was provided. Exiting.." in r.stderr


def test_help_and_unknown_options(cli):
    r = run(cli, "--i_ref=a.png", "--i_mov", "b.png", "-h")
    assert r.returncode == 0 and "Usage:" in r.stdout
    assert "Output image filename is output_image.png." in r.stderr
    assert "Fused Output image filename is output_image_fused.png." in r.stderr
    r = run(cli, "--i_ref", "a.png", "--i_mov", "b.png", "-ofoo.png", "-f", "bar.png", "-h")
    assert "Output image filename is foo.png." in r.stderr and "Fused Output image filename is bar.png." in r.stderr
    for bad in (["--nope", "1"], ["-z"], ["-m", "x1"], ["--i_ref"], ["--depth", "-1"]):
        r = run(cli, *bad)
        assert r.returncode == 1 and "cgs_matcher:" in r.stderr, bad


def test_unreadable_images(cli, tmp_path):
    bad = tmp_path / "bad.png"
    bad.write_bytes(b"not a png")
    r = run(cli, "--i_ref", str(tmp_path / "missing.png"), "--i_mov", str(bad), cwd=str(tmp_path))
    assert r.returncode == 1 and "Reference image " + str(tmp_path / "missing.png") + " failed to load!" in r.stderr
    from PIL import Image
    import numpy as np
    Image.fromarray(np.zeros((4, 4), np.uint8)).save(tmp_path / "ok.png")
    r = run(cli, "--i_ref", str(tmp_path / "ok.png"), "--i_mov", str(bad), cwd=str(tmp_path))
    assert r.returncode == 1 and "Moving image " + str(bad) + " failed to load!" in r.stderr
    assert not (tmp_path / "imageMatcherResults.txt").exists()
