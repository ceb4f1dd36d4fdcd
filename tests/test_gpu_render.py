"""GPU image parity: the gfx950 path tracer vs the oracle's restatement of
Aurora's shaders, on the same seeded scenes, cameras and sample indices.

Both evaluate the same PCG2D stream, draw order, BSDF and MIS; hits are
bit-identical (see test_gpu_trace.py), shading runs in FP32 on both sides with
different instruction rounding.  A path can therefore only diverge where a
float comparison (lobe choice, roulette) falls within rounding of its
threshold.  Tolerances:
  * per pixel: |gpu - cpu| <= 2e-3 * |cpu| + 2e-4 on RGBA for >= 99% of pixels;
  * image mean: |mean(gpu) - mean(cpu)| within 4 standard errors of the
    per-pixel difference (Monte Carlo confidence bound on the mean).
"""
import math

import numpy as np
import pytest

import pyoracle
from aurora_amd import capi, scenes
from aurora_amd.session import AphError

pytestmark = pytest.mark.gpu


def compare(gpu, cpu, min_frac=0.99):
    ok = np.isclose(gpu, cpu, rtol=2e-3, atol=2e-4).all(axis=-1)
    frac = ok.mean()
    d = (gpu - cpu)[..., :3].mean(axis=-1).ravel()
    se = d.std(ddof=1) / np.sqrt(d.size) if d.size > 1 else 0.0
    assert np.isfinite(gpu).all()
    assert frac >= min_frac, f"only {frac:.4f} of pixels match the oracle"
    assert abs(d.mean()) <= 4 * se + 1e-6, f"mean difference {d.mean()} exceeds 4 standard errors ({se})"
    return frac


def render_both(renderer, sc, spp, mode=capi.APH_MODE_MEGAKERNEL, start=0):
    tlas = renderer.load_scene(sc)
    tgt = renderer.target(sc.width, sc.height)
    renderer.render([renderer.request(tgt, tlas, sc.frame, start, spp)], mode)
    renderer.wait()
    gpu = renderer.readback(tgt, sc.width, sc.height)
    ora = pyoracle.OracleScene(sc)
    cpu, _ = ora.render(sc.frame, sc.width, sc.height, start, spp)
    return gpu, cpu


def test_teapot_fixture(renderer, teapot_path):
    sc = scenes.teapot_scene(teapot_path, 96, 96)
    gpu, cpu = render_both(renderer, sc, 4)
    compare(gpu, cpu)
    assert gpu[..., 3].mean() > 0.05  # the teapot is in view


def test_config1_scene(renderer):
    sc = scenes.config1(width=128, height=96, trace_depth=3)
    gpu, cpu = render_both(renderer, sc, 4)
    compare(gpu, cpu)


def test_config2_materials_transmission_coat_sun(renderer):
    sc = scenes.config2(subdiv=2, width=96, height=64, trace_depth=5, env_w=256, env_h=128)
    gpu, cpu = render_both(renderer, sc, 2)
    compare(gpu, cpu, min_frac=0.98)


def test_distant_lights_and_bilinear_env(renderer, teapot_path):
    sc = scenes.teapot_scene(teapot_path, 64, 64)
    sky = scenes.sky_gradient(64, 32, np.random.default_rng(5))
    sc.env = dict(light_rgb=sky, filter=1)
    fr = sc.frame
    for k, (d, col, inten) in enumerate([((0.3, 1.0, -0.5), (1, 0.9, 0.8), 2.0), ((-1.0, 0.5, 0.2), (0.3, 0.4, 1.0), 1.0)]):
        dd = np.array(d) / np.linalg.norm(d)
        for i in range(3):
            fr.lights[k].direction[i] = dd[i]
            fr.lights[k].color_intensity[i] = col[i]
        fr.lights[k].color_intensity[3] = inten
        fr.lights[k].cos_radius = np.cos(np.radians(1.0))
    for count in (1, 2):
        fr.light_count = count
        gpu, cpu = render_both(renderer, sc, 2)
        compare(gpu, cpu)


def test_progressive_accumulation_equals_single_call(renderer, teapot_path):
    """render(0,2) then render(2,2) == render(0,4): the running mean is keyed by sample index."""
    sc = scenes.teapot_scene(teapot_path, 48, 48)
    tlas = renderer.load_scene(sc)
    a = renderer.target(48, 48)
    b = renderer.target(48, 48)
    renderer.render([renderer.request(a, tlas, sc.frame, 0, 4)])
    renderer.render([renderer.request(b, tlas, sc.frame, 0, 2)])
    renderer.render([renderer.request(b, tlas, sc.frame, 2, 2)])
    renderer.wait()
    np.testing.assert_array_equal(renderer.readback(a, 48, 48), renderer.readback(b, 48, 48))


def test_post_processing_matches_oracle(renderer, teapot_path):
    sc = scenes.teapot_scene(teapot_path, 32, 32)
    tlas = renderer.load_scene(sc)
    tgt = renderer.target(32, 32)
    renderer.render([renderer.request(tgt, tlas, sc.frame, 0, 2)])
    acc = renderer.readback(tgt, 32, 32)
    for tone in (False, True):
        out = renderer.post(tgt, 32, 32, (1.2, 1.0, 0.9), tone_map=tone, gamma=True)
        ref = np.zeros_like(acc)
        import ctypes as C
        b = (C.c_float * 3)(1.2, 1.0, 0.9)
        pyoracle.load().ora_post(pyoracle._p(acc), 32 * 32, C.cast(b, C.c_void_p), int(tone), 1, pyoracle._p(ref))
        np.testing.assert_allclose(out, ref, rtol=1e-5, atol=1e-6)


# ---------------------------------------------------------------- wavefront pipeline

def test_wavefront_matches_oracle_and_megakernel(renderer, teapot_path):
    for sc, spp in ((scenes.teapot_scene(teapot_path, 64, 64), 4),
                    (scenes.config2(subdiv=2, width=96, height=64, trace_depth=5, env_w=256, env_h=128), 2)):
        gpu_wf, cpu = render_both(renderer, sc, spp, mode=capi.APH_MODE_WAVEFRONT)
        compare(gpu_wf, cpu, min_frac=0.98)
        gpu_mk, _ = render_both(renderer, sc, spp, mode=capi.APH_MODE_MEGAKERNEL)
        same = np.isclose(gpu_wf, gpu_mk, rtol=1e-5, atol=1e-6).all(axis=-1).mean()
        assert same >= 0.995, same


def test_wavefront_request_alone_equals_request_in_a_batch(renderer, teapot_path):
    """Serving determinism: a request's image is bit-identical whether it runs alone or
    shares the wavefront queues with other requests (different scenes, sizes, spp)."""
    a = scenes.teapot_scene(teapot_path, 40, 32)
    b = scenes.config1(width=56, height=48, trace_depth=3)
    ta = renderer.load_scene(a)
    tb = renderer.load_scene(b)
    alone = renderer.target(40, 32)
    renderer.render([renderer.request(alone, ta, a.frame, 0, 3)], capi.APH_MODE_WAVEFRONT)
    mixa = renderer.target(40, 32)
    mixb = renderer.target(56, 48)
    renderer.render([renderer.request(mixb, tb, b.frame, 0, 5), renderer.request(mixa, ta, a.frame, 0, 3)],
                    capi.APH_MODE_WAVEFRONT)
    renderer.wait()
    np.testing.assert_array_equal(renderer.readback(alone, 40, 32), renderer.readback(mixa, 40, 32))
    solo_b = renderer.target(56, 48)
    renderer.render([renderer.request(solo_b, tb, b.frame, 0, 5)], capi.APH_MODE_WAVEFRONT)
    np.testing.assert_array_equal(renderer.readback(solo_b, 56, 48), renderer.readback(mixb, 56, 48))


def test_wavefront_progressive_accumulation(renderer, teapot_path):
    sc = scenes.teapot_scene(teapot_path, 48, 40)
    tlas = renderer.load_scene(sc)
    a = renderer.target(48, 40)
    b = renderer.target(48, 40)
    renderer.render([renderer.request(a, tlas, sc.frame, 0, 5)], capi.APH_MODE_WAVEFRONT)
    renderer.render([renderer.request(b, tlas, sc.frame, 0, 2)], capi.APH_MODE_WAVEFRONT)
    renderer.render([renderer.request(b, tlas, sc.frame, 2, 3)], capi.APH_MODE_WAVEFRONT)
    renderer.wait()
    np.testing.assert_array_equal(renderer.readback(a, 48, 40), renderer.readback(b, 48, 40))


def test_nan_sample_direction_pixels_match_oracle(renderer):
    """config2 at 1920x1080: pixels (104, 906) and (823, 480) each draw a sample of
    exactly 1.0 within the first 24 sample indices, so sampleHemisphere returns a NaN
    direction that reaches the environment lookup (defined to read texel (0, 0)) and the
    NaN contribution is cleared by adjustRadiance.  Both paths must agree there."""
    sc = scenes.config2(subdiv=4)
    tlas = renderer.load_scene(sc)
    w, h, spp = sc.width, sc.height, 24
    pix = np.array([906 * w + 104, 480 * w + 823, 0, w * h - 1, 540 * w + 960], dtype=np.uint32)
    ora = pyoracle.OracleScene(sc)
    cpu, _ = ora.render(sc.frame, w, h, 0, spp, pixels=pix)
    for mode in (capi.APH_MODE_WAVEFRONT, capi.APH_MODE_MEGAKERNEL):
        tgt = renderer.target(w, h)
        renderer.render([renderer.request(tgt, tlas, sc.frame, 0, spp)], mode)
        renderer.wait()
        gpu = renderer.readback(tgt, w, h)
        g = gpu.reshape(-1, 4)[pix]
        c = cpu.reshape(-1, 4)[pix]
        assert np.isfinite(g).all()
        np.testing.assert_allclose(g, c, rtol=2e-3, atol=2e-4)


def test_material_layers_match_oracle(renderer, teapot_path):
    """Layers on the GPU: an opaque top layer equals the swapped base material bit for bit
    (same draws); a partly opaque layer matches the oracle's layer loop."""
    sc = scenes.teapot_scene(teapot_path, 48, 40)
    top = capi.default_material()
    top.base_color[:] = [0.1, 0.8, 0.2]
    top.metalness = 1.0
    sc.materials = [sc.materials[0], top]
    sc.layers = [[1]]
    for mode in (capi.APH_MODE_MEGAKERNEL, capi.APH_MODE_WAVEFRONT):
        gl, _ = render_both(renderer, sc, 3, mode=mode)
        sw = scenes.teapot_scene(teapot_path, 48, 40)
        sw.materials = [top]
        gs, _ = render_both(renderer, sw, 3, mode=mode)
        np.testing.assert_array_equal(gl, gs)
    half = capi.default_material()
    half.base_color[:] = [0.9, 0.3, 0.1]
    half.opacity[:] = [0.5, 0.5, 0.5]
    sc.materials = [sc.materials[0], half]
    for mode in (capi.APH_MODE_MEGAKERNEL, capi.APH_MODE_WAVEFRONT):
        gpu, cpu = render_both(renderer, sc, 4, mode=mode)
        compare(gpu, cpu)


def test_graph_replays_equal_eager_launches(renderer):
    """Each chunk's stage sequence is one captured HIP graph.  With 64K-path chunks a
    96x64x16 render replays it twice per call; every replay - with eager work and
    synchronizations between them - must give the eager image bit for bit, and so
    must a different launch configuration (tuning never changes results)."""
    sc = scenes.config2(subdiv=2, width=96, height=64, trace_depth=5, env_w=256, env_h=128)
    tlas = renderer.load_scene(sc)
    renderer.set_tuning(1 << 16, 16, 8)

    def render(spp=16, start=0, tgt=None):
        tgt = tgt or renderer.target(96, 64)
        renderer.render([renderer.request(tgt, tlas, sc.frame, start, spp)], capi.APH_MODE_WAVEFRONT)
        renderer.wait()
        return tgt, renderer.readback(tgt, 96, 64)

    renderer.set_launch_mode(capi.APH_LAUNCH_EAGER)
    _, eager = render()
    renderer.set_launch_mode(capi.APH_LAUNCH_GRAPH)
    for k in range(3):
        _, g = render()
        np.testing.assert_array_equal(g, eager)
        other = renderer.target(96, 64)  # eager work between replays
        renderer.render([renderer.request(other, tlas, sc.frame, 0, 1)], capi.APH_MODE_MEGAKERNEL)
        renderer.wait()
    t, _ = render(8, 0)
    _, prog = render(8, 8, t)
    np.testing.assert_array_equal(prog, eager)
    renderer.set_tuning(1 << 17, 8, 16)
    _, tuned = render()
    np.testing.assert_array_equal(tuned, eager)
    for refill, wait in ((1, 0), (64, 0), (16, 63)):  # refill threshold; lanes still searching at leaf time
        renderer.set_tuning(1 << 17, 8, 16, refill, wait)
        _, tuned = render()
        np.testing.assert_array_equal(tuned, eager)
    renderer.set_tuning(4 << 20, 16, 8)


def test_scenes_keep_their_own_environments(renderer):
    """Two resident scenes with different lat-long environments, each set on its TLAS
    (aph_environment_create / aph_tlas_set_environment), rendered together in one call:
    each image equals that scene rendered alone with its environment as the context's,
    bit for bit, in both modes.  An environment set on a TLAS cannot be destroyed."""
    r = renderer
    w, h = 64, 48
    a = scenes.config1(seed=31, width=w, height=h, trace_depth=3)
    b = scenes.config1(seed=32, width=w, height=h, trace_depth=3)
    assert not np.array_equal(a.env["light_rgb"], b.env["light_rgb"])
    ta, tb = r.load_scenes([a, b], capi.APH_BUILDER_SAH)
    mixed = {}
    for mode in (capi.APH_MODE_WAVEFRONT, capi.APH_MODE_MEGAKERNEL):
        ga, gb = r.target(w, h), r.target(w, h)
        r.render([r.request(ga, ta, a.frame, 0, 8), r.request(gb, tb, b.frame, 0, 8)], mode)
        r.wait()
        mixed[mode] = (r.readback(ga, w, h), r.readback(gb, w, h))
    assert not np.array_equal(mixed[capi.APH_MODE_WAVEFRONT][0], mixed[capi.APH_MODE_WAVEFRONT][1])
    e = r.create_environment(b.env)
    r.set_environment(ta, e)
    with pytest.raises(AphError):
        r.destroy_environment(e)
    r.set_environment(ta, 0)
    r.destroy_environment(e)
    for k, sc in enumerate((a, b)):
        t = r.load_scene(sc, capi.APH_BUILDER_SAH)  # the scene's environment as the context's
        for mode, imgs in mixed.items():
            g = r.target(w, h)
            r.render([r.request(g, t, sc.frame, 0, 8)], mode)
            r.wait()
            np.testing.assert_array_equal(r.readback(g, w, h), imgs[k])


def test_autotune_keeps_results(renderer, teapot_path):
    sc = scenes.teapot_scene(teapot_path, 64, 48)
    tlas = renderer.load_scene(sc)
    ref = renderer.target(64, 48)
    renderer.render([renderer.request(ref, tlas, sc.frame, 0, 4)], capi.APH_MODE_WAVEFRONT)
    renderer.wait()
    before = renderer.readback(ref, 64, 48)
    tgt = renderer.target(64, 48)
    untouched = renderer.readback(tgt, 64, 48)
    best = renderer.autotune([renderer.request(tgt, tlas, sc.frame, 4, 4)])
    # trials render into scratch targets: the request's own target is left as it was
    np.testing.assert_array_equal(renderer.readback(tgt, 64, 48), untouched)
    # the candidate set of aph_wavefront_autotune
    assert best.ms > 0 and best.chunk_paths in (4 << 20, 8 << 20, 16 << 20, 32 << 20, 64 << 20)
    assert best.trace_blocks_per_cu in (4, 8, 16) and best.shade_blocks_per_cu in (4, 8, 16)
    assert best.trace_refill_lanes in (16, 24, 32, 40, 48, 56, 64)
    assert best.trace_leaf_wait_lanes in (0, 8, 16, 24, 32)
    t = renderer.tuning()
    assert (t.chunk_paths, t.trace_blocks_per_cu, t.shade_blocks_per_cu, t.trace_refill_lanes,
            t.trace_leaf_wait_lanes) == (best.chunk_paths, best.trace_blocks_per_cu, best.shade_blocks_per_cu,
                                         best.trace_refill_lanes, best.trace_leaf_wait_lanes)
    after = renderer.target(64, 48)
    renderer.render([renderer.request(after, tlas, sc.frame, 0, 4)], capi.APH_MODE_WAVEFRONT)
    renderer.wait()
    np.testing.assert_array_equal(renderer.readback(after, 64, 48), before)
    renderer.set_tuning(4 << 20, 16, 8)


def test_textured_scene_matches_oracle(renderer):
    """Every texture slot, UV transform, address mode and image format (scenes.textured_scene):
    the GPU's sampling (device sinf/cosf in the UV rotation) against the oracle's, same bar
    as the untextured scenes."""
    sc = scenes.textured_scene(width=96, height=64)
    for mode in (capi.APH_MODE_MEGAKERNEL, capi.APH_MODE_WAVEFRONT):
        gpu, cpu = render_both(renderer, sc, 4, mode=mode)
        compare(gpu, cpu)


@pytest.mark.parametrize("slot, value", [
    (capi.APH_TEX_BASE_COLOR, (0.7, 0.2, 0.4, 1.0)),
    (capi.APH_TEX_SPECULAR_ROUGHNESS, (0.35, 0.0, 0.0, 1.0)),
    (capi.APH_TEX_EMISSION_COLOR, (0.9, 0.6, 0.1, 1.0)),
    (capi.APH_TEX_OPACITY, (0.6, 0.6, 0.6, 1.0)),
])
def test_constant_texture_equals_constant_material(renderer, slot, value):
    """A constant image in a slot renders bit-identically to the material constant it replaces."""
    from test_oracle_textures import constant_scene

    tex, plain = constant_scene(slot, value)
    for mode in (capi.APH_MODE_MEGAKERNEL, capi.APH_MODE_WAVEFRONT):
        images = []
        for sc in (tex, plain):
            tlas = renderer.load_scene(sc)
            tgt = renderer.target(sc.width, sc.height)
            renderer.render([renderer.request(tgt, tlas, sc.frame, 0, 2)], mode)
            renderer.wait()
            images.append(renderer.readback(tgt, sc.width, sc.height))
        np.testing.assert_array_equal(images[0], images[1])


def test_textured_guides_match_oracle(renderer):
    """Guide images of a textured scene: base colour and roughness from the textures, the
    normal from the normal map.  The UV rotation uses the device's and the host's
    sinf/cosf; where the bilinear filter crosses a sharp texture edge (the checker) a
    last-bit difference of the coordinate becomes up to (texels per unit) x (edge
    contrast) x (coordinate ulp).  Bar: >= 99.9 % of values within 1e-4 relative +
    1e-5 absolute, all within 1e-3 absolute."""
    w, h = 64, 40
    sc = scenes.textured_scene(width=w, height=h)
    tlas = renderer.load_scene(sc)
    fr = sc.frame
    vp = scenes.view_proj(fr)
    ivp = np.array(fr.inv_view, dtype=np.float32)
    aov = renderer.aov(w, h)
    renderer.render_aovs(aov, tlas, fr, vp, vp, ivp, sample_index=2)
    gd, gn, gb, _ = renderer.aov_readback(aov, w, h)
    od, on, ob, _ = pyoracle.OracleScene(sc).aovs(fr, w, h, 2, vp, vp, ivp, -2000.0)
    np.testing.assert_allclose(gd, od, rtol=1e-5)
    for g, o in ((gb, ob), (gn, on)):
        assert np.isclose(g, o, rtol=1e-4, atol=1e-5).mean() >= 0.999
        assert np.abs(g - o).max() <= 1e-3
    assert np.unique(np.round(gb[..., :3], 3).reshape(-1, 3), axis=0).shape[0] > 20  # textured, not constant


def test_texture_errors(renderer):
    from aurora_amd.session import AphError

    r = renderer
    sc = scenes.textured_scene(width=16, height=16)
    tlas = r.load_scene(sc)
    tgt = r.target(16, 16)
    img = r.image(np.zeros((4, 4, 4), np.uint8))
    n = len(sc.materials)
    with pytest.raises(AphError) as e:  # a flagged slot loses its image
        r.bind_textures(n, {(0, capi.APH_TEX_BASE_COLOR): (img, 0, 0)})
        r.render([r.request(tgt, tlas, sc.frame, 0, 1)])
    assert e.value.code == capi.APH_E_STATE
    with pytest.raises(AphError) as e:
        r.bind_textures(n, {(0, capi.APH_TEX_BASE_COLOR): (img, 7, 0)})
    assert e.value.code == capi.APH_E_INVALID
    with pytest.raises(AphError) as e:
        r.bind_textures(n - 1, {})
    assert e.value.code == capi.APH_E_INVALID
    r.load_scene(sc)  # the full bindings again
    bound = img
    r.bind_textures(n, {**{k: v for k, v in
                           r._scene_textures(sc).items()}, (1, capi.APH_TEX_NORMAL): (bound, 0, 0)})
    with pytest.raises(AphError) as e:
        r.destroy_image(bound)
    assert e.value.code == capi.APH_E_STATE
    r.materials(sc.materials)  # uploading clears the bindings
    r.destroy_image(bound)
    d = capi.ImageDesc(4, 4, 9, 0, np.zeros(64, np.uint8).ctypes.data)
    import ctypes as C
    h = C.c_uint64()
    assert r.lib.aph_image_create(r.ctx, C.byref(d), C.byref(h)) == capi.APH_E_INVALID


def test_duplicate_targets_are_rejected_before_any_launch(renderer, teapot_path):
    """The wavefront folds a chunk's samples into each target in one pass, so two
    requests of one call may not share a target; a bad request in either mode leaves
    every target of the call untouched (all requests are checked before any launch)."""
    r = renderer
    sc = scenes.teapot_scene(teapot_path, 16, 16)
    tlas = r.load_scene(sc)
    t = r.target(16, 16)
    before = r.readback(t, 16, 16)
    with pytest.raises(AphError) as e:
        r.render([r.request(t, tlas, sc.frame, 0, 2), r.request(t, tlas, sc.frame, 2, 2)], capi.APH_MODE_WAVEFRONT)
    assert e.value.code == capi.APH_E_INVALID
    for mode in (capi.APH_MODE_WAVEFRONT, capi.APH_MODE_MEGAKERNEL):
        with pytest.raises(AphError) as e:
            r.render([r.request(t, tlas, sc.frame, 0, 2), r.request(t + 10 ** 6, tlas, sc.frame, 0, 2)], mode)
        assert e.value.code == capi.APH_E_INVALID
    r.wait()
    np.testing.assert_array_equal(r.readback(t, 16, 16), before)
    # megakernel requests on one target run in order: the same as one progressive call
    u = r.target(16, 16)
    r.render([r.request(u, tlas, sc.frame, 0, 2), r.request(u, tlas, sc.frame, 2, 2)], capi.APH_MODE_MEGAKERNEL)
    v = r.target(16, 16)
    r.render([r.request(v, tlas, sc.frame, 0, 4)], capi.APH_MODE_MEGAKERNEL)
    r.wait()
    np.testing.assert_array_equal(r.readback(u, 16, 16), r.readback(v, 16, 16))


def test_production_chunk_shape_equals_small_chunks(renderer):
    """The autotuned launch shape of the headline frame: 64M-path chunks (about 35 GB of
    path state and queues), 1920x1080 at 33 spp = 68.4M paths, so one call runs two
    chunks of the captured graph back to back.  The image must equal, bit for bit, the
    same render cut into 4M-path chunks (17 chunks), and the eager launches."""
    sc = scenes.config2(subdiv=4)
    tlas = renderer.load_scene(sc, capi.APH_BUILDER_SAH)
    w, h, spp = sc.width, sc.height, 33
    imgs = []
    for chunk, tr, sh, mode in ((64 << 20, 16, 8, capi.APH_LAUNCH_GRAPH), (4 << 20, 16, 8, capi.APH_LAUNCH_GRAPH),
                                (64 << 20, 16, 8, capi.APH_LAUNCH_EAGER)):
        renderer.set_tuning(chunk, tr, sh, 16, 16)
        renderer.set_launch_mode(mode)
        tgt = renderer.target(w, h)
        renderer.render([renderer.request(tgt, tlas, sc.frame, 0, spp)], capi.APH_MODE_WAVEFRONT)
        renderer.wait()
        imgs.append(renderer.readback(tgt, w, h))
    renderer.set_launch_mode(capi.APH_LAUNCH_GRAPH)
    renderer.set_tuning(4 << 20, 16, 8)
    assert np.isfinite(imgs[0]).all() and imgs[0][..., 3].mean() > 0.05
    np.testing.assert_array_equal(imgs[0], imgs[1])
    np.testing.assert_array_equal(imgs[0], imgs[2])


# ---------------------------------------------------------------- every lobe, camera and environment mode

def near_delta_mask(sc, instances, grow=2):
    """Pixels whose pixel-centre camera ray hits one of `instances`, grown by `grow` pixels."""
    import ctypes as C
    W, H = sc.width, sc.height
    fr = scenes.look_at_frame((0.0, 0.0, -5.0), (0.0, 0.0, 0.0), (0, 1, 0), 45.0, W / H, 1)
    fr.temporal_jitter = 1
    rays = np.zeros(W * H, dtype=capi.RAY_DTYPE)
    lib = pyoracle.load()
    o, d = np.zeros(3, np.float32), np.zeros(3, np.float32)
    for y in range(H):
        for x in range(W):
            lib.ora_camera_ray(C.byref(fr), W, H, x, y, 0, pyoracle._p(o), pyoracle._p(d))
            rays[y * W + x]["org"], rays[y * W + x]["dir"] = o, d
    rays["tmax"] = np.inf
    hit = np.isin(pyoracle.OracleScene(sc).trace_closest(rays)["inst"], instances).reshape(H, W)
    out = hit.copy()
    for dy in range(-grow, grow + 1):
        for dx in range(-grow, grow + 1):
            out |= np.roll(np.roll(hit, dy, 0), dx, 1)
    return out


@pytest.mark.parametrize("kind", scenes.SWEEP_KINDS)
def test_material_sweep_matches_oracle(renderer, teapot_path, kind):
    """The 5 x 5 teapot sweeps of TestMaterialAdvancedMaterialProperties (TestMaterial.cpp:613):
    subsurface, sheen, coat, coat_affect_*, transmission, plus anisotropy + rotation (main and
    coat) and thin-walled transmission, in both modes, against the oracle.
    The coat sweep's first row has coat_roughness 0, which mx_roughness_anisotropy clamps to
    alpha = FLT_EPS: a numerical delta whose NDF (~1 / alpha^2) turns last-bit differences of
    the half vector into large value differences (measured: 40 % of those teapots' pixels
    differ beyond the per-pixel bar, every other pixel of every sweep matches).  Those five
    teapots (and a 2-pixel border) are held to the image-mean bound only."""
    sc = scenes.material_sweep(teapot_path, kind, 128, 128)
    keep = ~near_delta_mask(sc, [0, 1, 2, 3, 4]) if kind == "coat" else np.ones((128, 128), bool)
    for mode in (capi.APH_MODE_MEGAKERNEL, capi.APH_MODE_WAVEFRONT):
        gpu, cpu = render_both(renderer, sc, 8, mode=mode)
        compare(gpu[keep][None], cpu[keep][None])
        compare(gpu, cpu, min_frac=0.0)  # the mean bound over the whole image


def test_headline_frame_random_pixels_match_oracle(renderer):
    """The benchmark frame itself (config2 at subdivision 4: 1.02M instanced triangles,
    1920x1080, 2048x1024 sun sky, transmission, coat, a decal layer): 2,048 random pixels
    at 8 spp against the oracle's render of the same pixels, in both modes."""
    sc = scenes.config2(subdiv=4)
    tlas = renderer.load_scene(sc, capi.APH_BUILDER_SAH)
    w, h, spp = sc.width, sc.height, 8
    pix = np.sort(np.random.default_rng(42).choice(w * h, 2048, replace=False)).astype(np.uint32)
    cpu, _ = pyoracle.OracleScene(sc).render(sc.frame, w, h, 0, spp, pixels=pix)
    c = cpu.reshape(-1, 4)[pix]
    for mode in (capi.APH_MODE_WAVEFRONT, capi.APH_MODE_MEGAKERNEL):
        tgt = renderer.target(w, h)
        renderer.render([renderer.request(tgt, tlas, sc.frame, 0, spp)], mode)
        renderer.wait()
        g = renderer.readback(tgt, w, h).reshape(-1, 4)[pix]
        compare(g[None], c[None])
        assert (c[:, 3] > 0).mean() > 0.3  # most sampled pixels see geometry


def test_orthographic_and_thin_lens_cameras_match_oracle(renderer, teapot_path):
    """computeCameraRay's orthographic variant and depth of field (RayTrace.slang:130-182,
    the lens taking the draw after the jitter), rendered as the reference's camera test
    does (TestRenderer.cpp:1124): GPU against the oracle in both modes."""
    ortho = scenes.teapot_scene(teapot_path, 64, 64)
    ortho.frame.is_ortho = 1
    ortho.frame.view_size[0] = ortho.frame.view_size[1] = 4.5
    lens = scenes.teapot_scene(teapot_path, 64, 64)
    lens.frame.lens_radius = 0.25
    lens.frame.focal_distance = 4.0
    for sc in (ortho, lens):
        for mode in (capi.APH_MODE_MEGAKERNEL, capi.APH_MODE_WAVEFRONT):
            gpu, cpu = render_both(renderer, sc, 8, mode=mode)
            compare(gpu, cpu)
    assert render_both(renderer, lens, 8)[0][..., 3].mean() > 0.05


def test_environment_transforms_and_screen_background_match_oracle(renderer, teapot_path):
    """EnvironmentBase's light and background transforms (EnvironmentBase.cpp:52-71) on a
    textured sky and background, and background_use_screen (Environment.slang:68-74): GPU
    against the oracle in both modes."""
    rng = np.random.default_rng(8)
    sky = scenes.sun_sky(256, 128, rng, sun_radiance=200.0, sun_radius_deg=4.0)
    bg = scenes.sky_gradient(96, 48, np.random.default_rng(9))
    ry = [0.0] * 16
    c, s = math.cos(1.1), math.sin(1.1)
    ry[0], ry[2], ry[5], ry[8], ry[10], ry[15] = c, s, 1.0, -s, c, 1.0
    rx = [1.0, 0, 0, 0, 0, math.cos(0.4), -math.sin(0.4), 0, 0, math.sin(0.4), math.cos(0.4), 0, 0, 0, 0, 1.0]
    for env in (dict(light_rgb=sky, background_rgb=bg, light_transform=ry, background_transform=rx),
                dict(light_rgb=sky, background_rgb=bg, background_use_screen=1, filter=1)):
        sc = scenes.teapot_scene(teapot_path, 64, 48)
        sc.env = env
        for mode in (capi.APH_MODE_MEGAKERNEL, capi.APH_MODE_WAVEFRONT):
            gpu, cpu = render_both(renderer, sc, 4, mode=mode)
            compare(gpu, cpu)


def test_transparent_shadows_match_oracle(renderer, teapot_path):
    """Shadows through non-opaque surfaces (force_opaque_shadows = 0: the DirectX shadow
    any-hit shader, HitShaderEntryPoints.slang:489-553): the stacked-quad scene of the oracle
    pin, straight-down rays through the quads' shared diagonal included, and config2's
    transmissive and coated spheres under a distant light, in both modes against the oracle."""
    from test_oracle_pathtracer import shadow_stack_scene

    def mat(opacity, transmission=0.0, tcolor=(1, 1, 1)):
        m = capi.default_material()
        m.opacity[:] = opacity
        m.transmission = transmission
        m.transmission_color[:] = tcolor
        return m

    stack = shadow_stack_scene([mat((0.3, 0.5, 0.7)), mat((1.0, 1.0, 1.0), 0.5, (1.0, 0.5, 0.0))])
    stack.width = stack.height = 32
    fr = scenes.look_at_frame((0.0, 0.5, 0.0), (0.0, 0.0, 0.0), (0, 0, -1), 2.0, 1.0, 1)
    fr.light_count = 1
    fr.lights[0].direction[:] = [0.0, 1.0, 0.0]
    fr.lights[0].color_intensity[:] = [1.0, 0.9, 0.8, 3.0]
    fr.lights[0].cos_radius = 1.0
    fr.force_opaque_shadows = 0
    stack.frame = fr
    c2 = scenes.config2(subdiv=2, width=96, height=64, trace_depth=5, env_w=256, env_h=128)
    c2.frame.force_opaque_shadows = 0
    c2.frame.light_count = 1
    d = np.array([0.3, 0.8, -0.4]) / np.linalg.norm([0.3, 0.8, -0.4])
    c2.frame.lights[0].direction[:] = d
    c2.frame.lights[0].color_intensity[:] = [1.0, 0.95, 0.9, 2.0]
    c2.frame.lights[0].cos_radius = math.cos(math.radians(2.0))
    for sc, spp, bar in ((stack, 4, 1.0), (c2, 2, 0.98)):
        for mode in (capi.APH_MODE_MEGAKERNEL, capi.APH_MODE_WAVEFRONT):
            gpu, cpu = render_both(renderer, sc, spp, mode=mode)
            compare(gpu, cpu, min_frac=bar)
