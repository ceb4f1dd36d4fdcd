"""aph_render_noalloc, the allocation-free render: its images equal aph_render's wavefront
images bit for bit (same pipeline, chunked to the fixed array's 131072 paths), on two calls
in a row with the second smaller, and calls beyond its limits are refused."""
import ctypes as C

import numpy as np
import pytest

from aurora_amd import capi, scenes
from aurora_amd.session import AphError

pytestmark = pytest.mark.gpu


def render_noalloc(r, reqs):
    arr = (capi.Request * len(reqs))(*reqs)
    r._check(r.lib.aph_render_noalloc(r.ctx, C.cast(arr, C.c_void_p), len(reqs)))


def test_noalloc_render_equals_wavefront_render(renderer, teapot_path):
    r = renderer
    a = scenes.teapot_scene(teapot_path, 320, 240)  # 76800 pixels x 6 spp: four chunks of at most 131072 paths
    b = scenes.config1(width=96, height=64, trace_depth=3)
    ta, tb = r.load_scenes([a, b], capi.APH_BUILDER_SAH)
    calls = [[(ta, a.frame, 320, 240, 0, 6), (tb, b.frame, 96, 64, 0, 9)],  # two requests, several chunks
             [(tb, b.frame, 96, 64, 3, 2)]]                                 # then a smaller call
    for call in calls:
        ref = [r.target(w, h) for _, _, w, h, _, _ in call]
        mine = [r.target(w, h) for _, _, w, h, _, _ in call]
        r.render([r.request(t, tl, fr, s0, n) for t, (tl, fr, w, h, s0, n) in zip(ref, call)],
                 capi.APH_MODE_WAVEFRONT)
        render_noalloc(r, [r.request(t, tl, fr, s0, n) for t, (tl, fr, w, h, s0, n) in zip(mine, call)])
        r.wait()
        for t0, t1, (_, _, w, h, _, _) in zip(ref, mine, call):
            img = r.readback(t1, w, h)
            np.testing.assert_array_equal(img, r.readback(t0, w, h))
            assert np.isfinite(img).all() and img[..., 3].mean() > 0.01


def test_noalloc_render_refuses_what_does_not_fit(renderer, teapot_path):
    r = renderer
    sc = scenes.teapot_scene(teapot_path, 512, 512)  # 262144 pixels: above the 131072 paths in flight
    tlas = r.load_scene(sc)
    big = r.target(512, 512)
    with pytest.raises(AphError) as e:
        render_noalloc(r, [r.request(big, tlas, sc.frame, 0, 1)])
    assert e.value.code == capi.APH_E_INVALID
    small = r.target(8, 8)
    with pytest.raises(AphError) as e:  # 100,000 one-sample segments: beyond the 1 MiB call table
        render_noalloc(r, [r.request(small, tlas, sc.frame, 0, 100000)])
    assert e.value.code == capi.APH_E_INVALID
    np.testing.assert_array_equal(r.readback(big, 512, 512), 0.0)
