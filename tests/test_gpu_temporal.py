"""GPU guide images (aph_render_aovs) and temporal resolve (aph_temporal_resolve)
against the oracle's restatements on the same inputs.

Bar: AOVs - hit/miss pattern and base colour exact, view depth and normals
within 1e-5 relative, motion within 1e-3 pixel (the camera ray and hit are
bit-identical; the remaining float arithmetic rounds alike up to contraction);
temporal resolve within 1e-5 relative + 1e-6 absolute on >= 99.9 % of pixels
(a rejection test can flip where a quantity rounds across its threshold).
"""
import numpy as np
import pytest

import pyoracle
from aurora_amd import capi, scenes

pytestmark = pytest.mark.gpu


def moving_cameras(w, h, frame_index=3):
    cur = scenes.look_at_frame((0.0, 6.0, 14.0), (0.0, 0.5, 0.0), (0, 1, 0), 45.0, w / h, 3)
    prev = scenes.look_at_frame((0.6, 6.2, 14.3), (0.1, 0.5, 0.0), (0, 1, 0), 45.0, w / h, 3)
    jit = np.zeros(2, dtype=np.float32)
    pyoracle.load().ora_halton_jitter(frame_index, 0, pyoracle._p(jit))
    cur.temporal_jitter = 1
    cur.jitter[:] = [float(jit[0]), float(jit[1])]
    return cur, prev


def test_aovs_match_oracle(renderer):
    sc = scenes.config1(width=48, height=32, trace_depth=3)
    tlas = renderer.load_scene(sc)
    w, h = 48, 32
    cur, prev = moving_cameras(w, h)
    vp, vpp = scenes.view_proj(cur), scenes.view_proj(prev)
    ivp = np.array(prev.inv_view, dtype=np.float32)
    aov = renderer.aov(w, h)
    renderer.render_aovs(aov, tlas, cur, vp, vpp, ivp, sky_view_z=-2000.0, sample_index=5)
    gd, gn, gb, gm = renderer.aov_readback(aov, w, h)
    od, on, ob, om = pyoracle.OracleScene(sc).aovs(cur, w, h, 5, vp, vpp, ivp, -2000.0)
    hit = od != -2000.0
    assert hit.mean() > 0.3 and (~hit).sum() > 0
    np.testing.assert_array_equal(gd == -2000.0, ~hit)
    np.testing.assert_array_equal(gb, ob)
    np.testing.assert_allclose(gd, od, rtol=1e-5)
    np.testing.assert_allclose(gn, on, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(gm, om, rtol=1e-4, atol=1e-3)
    assert np.abs(gm[hit][:, :2]).max() > 0.5  # the camera moved


@pytest.mark.parametrize("reset", [False, True])
def test_temporal_resolve_matches_oracle(renderer, reset):
    sc = scenes.config1(width=64, height=40, trace_depth=3)
    tlas = renderer.load_scene(sc)
    w, h = 64, 40
    cur_cam, prev_cam = moving_cameras(w, h)
    aov = renderer.aov(w, h)
    renderer.render_aovs(aov, tlas, cur_cam, scenes.view_proj(cur_cam), scenes.view_proj(prev_cam),
                         np.array(prev_cam.inv_view, dtype=np.float32))
    cur = renderer.target(w, h)
    renderer.render([renderer.request(cur, tlas, cur_cam, 0, 2)], capi.APH_MODE_WAVEFRONT)
    hin = renderer.target(w, h)
    renderer.render([renderer.request(hin, tlas, prev_cam, 0, 8)], capi.APH_MODE_WAVEFRONT)
    renderer.wait()
    hist = renderer.readback(hin, w, h)
    hist[3, 5, :3] = np.nan  # a poisoned history texel must not propagate
    renderer.target_upload(hin, hist)
    hout = renderer.target(w, h)
    renderer.temporal_resolve(cur, aov, hin, hout, is_reset=reset)
    renderer.wait()
    g = renderer.readback(hout, w, h)
    d, _, _, m = renderer.aov_readback(aov, w, h)
    o = pyoracle.temporal_resolve(renderer.readback(cur, w, h), d, m, hist, is_reset=int(reset))
    assert np.isfinite(g).all()
    ok = np.isclose(g, o, rtol=1e-5, atol=1e-6).all(axis=-1)
    assert ok.mean() >= 0.999, ok.mean()
    if reset:
        # alpha = 1: history + (current - history) rounds to current within an ulp or two
        np.testing.assert_allclose(g, renderer.readback(cur, w, h), rtol=1e-6, atol=1e-7)


def test_static_jittered_sequence_converges(renderer):
    """A static camera with Halton jitter, 1 spp per frame, resolved over 24 frames, gets closer to
    a 256-spp image than any single frame."""
    sc = scenes.config1(width=48, height=32, trace_depth=3)
    tlas = renderer.load_scene(sc)
    w, h = 48, 32
    ref = renderer.target(w, h)
    renderer.render([renderer.request(ref, tlas, sc.frame, 0, 256)], capi.APH_MODE_WAVEFRONT)
    aov = renderer.aov(w, h)
    hist = [renderer.target(w, h), renderer.target(w, h)]
    cur = renderer.target(w, h)
    vp = scenes.view_proj(sc.frame)
    jit = np.zeros(2, dtype=np.float32)
    first = None
    for k in range(24):
        fr = sc.frame
        pyoracle.load().ora_halton_jitter(k, 0, pyoracle._p(jit))
        fr.temporal_jitter = 1
        fr.jitter[:] = [float(jit[0]), float(jit[1])]
        renderer.render([renderer.request(cur, tlas, fr, k, 1)], capi.APH_MODE_WAVEFRONT)
        renderer.render_aovs(aov, tlas, fr, vp, vp, np.array(fr.inv_view, dtype=np.float32), sample_index=k)
        renderer.temporal_resolve(cur, aov, hist[k % 2], hist[(k + 1) % 2], is_reset=(k == 0))
        if first is None:
            renderer.wait()
            first = renderer.readback(cur, w, h)
    renderer.wait()
    r = renderer.readback(ref, w, h)[..., :3]
    res = renderer.readback(hist[24 % 2], w, h)[..., :3]
    rmse = lambda a: float(np.sqrt(np.mean((a - r) ** 2)))  # noqa: E731
    assert rmse(res) < 0.6 * rmse(first[..., :3]), (rmse(res), rmse(first[..., :3]))
