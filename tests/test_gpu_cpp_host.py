"""A C++ host of the C-ABI (tests/cpp/fixture_flow.cpp, built with g++ against include/aph.h
alone) runs the reference fixture's flow - 128x128, render(start, 4) called 32 times with start
advancing (AuroraTestHelpers.cpp:826-852), the fixture camera (:617-622) on the fixture teapot -
and its image equals the same flow driven from Python bit for bit, and the oracle's render of
the same 128 samples per pixel to the render bar."""
import os
import subprocess

import numpy as np
import pytest

import pyoracle
from aurora_amd import capi, scenes
from test_gpu_render import compare

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "_build", "fixture_flow")


def test_cpp_host_runs_the_fixture_flow(renderer, teapot_path, tmp_path):
    assert os.path.exists(BIN), "tests/cpp/_build/fixture_flow is not built (run the build)"
    pos, nrm, idx = scenes.load_teapot(teapot_path)
    mesh = tmp_path / "teapot.bin"
    with open(mesh, "wb") as f:
        np.array([len(pos), len(idx)], np.uint32).tofile(f)
        np.ascontiguousarray(pos, np.float32).tofile(f)
        np.ascontiguousarray(nrm, np.float32).tofile(f)
        np.ascontiguousarray(idx, np.uint32).tofile(f)
    out = tmp_path / "image.rgba32f"
    p = subprocess.run([BIN, str(mesh), str(out), "128", "128", "32", "4"], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    img = np.fromfile(out, np.float32).reshape(128, 128, 4)

    sc = scenes.teapot_scene(teapot_path, 128, 128)
    tlas = renderer.load_scene(sc, capi.APH_BUILDER_SAH)
    tgt = renderer.target(128, 128)
    for k in range(32):
        renderer.render([renderer.request(tgt, tlas, sc.frame, 4 * k, 4)], capi.APH_MODE_WAVEFRONT)
    renderer.wait()
    np.testing.assert_array_equal(img, renderer.readback(tgt, 128, 128))
    cpu, _ = pyoracle.OracleScene(sc).render(sc.frame, 128, 128, 0, 128)
    compare(img, cpu)
    assert img[..., 3].mean() > 0.05
