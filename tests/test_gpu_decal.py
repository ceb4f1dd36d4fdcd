"""Textured decals: material layers with geometry of their own (LayerDefinition's geometry,
PTScene.cpp:1051-1066; computeLayerUV, GlobalBufferAccessors.slang:116-127) on the GPU
against the oracle in both render modes, the layer-coordinate invariants bit for bit, and
the entry point's errors."""
import numpy as np
import pytest

from aurora_amd import capi, scenes
from aurora_amd.session import AphError
from test_gpu_render import compare, render_both

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", [capi.APH_MODE_MEGAKERNEL, capi.APH_MODE_WAVEFRONT])
def test_decal_matches_oracle(renderer, teapot_path, mode):
    sc = scenes.decal_scene(teapot_path)
    gpu, cpu = render_both(renderer, sc, 4, mode=mode)
    compare(gpu, cpu)
    plain = scenes.decal_scene(teapot_path)
    plain.layer_geometry = None
    other, _ = render_both(renderer, plain, 4, mode=mode)
    assert np.abs(gpu - other)[..., :3].mean() > 0.005  # the decal moved


def image(renderer, sc, mode=capi.APH_MODE_WAVEFRONT):
    tl = renderer.load_scene(sc)
    tgt = renderer.target(sc.width, sc.height)
    renderer.render([renderer.request(tgt, tl, sc.frame, 0, 2)], mode)
    renderer.wait()
    out = renderer.readback(tgt, sc.width, sc.height)
    renderer.destroy_target(tgt)
    return out, tl


def test_layer_geometry_equals_coordinates_on_the_instance(renderer, teapot_path):
    a = scenes.decal_scene(teapot_path, 64, 64)
    a.materials = [scenes._material((0.8, 0.8, 0.8), 0.0, 0.3), a.materials[1]]
    a.textures = {k: v for k, v in a.textures.items() if k[0] != 0}
    b = scenes.decal_scene(teapot_path, 64, 64)
    b.materials, b.textures = a.materials, a.textures
    b.meshes = [b.meshes[0]]
    b.uvs = {0: a.uvs[1]}
    b.layer_geometry = None
    for mode in (capi.APH_MODE_MEGAKERNEL, capi.APH_MODE_WAVEFRONT):
        assert np.array_equal(image(renderer, a, mode)[0], image(renderer, b, mode)[0])


def test_layer_geometry_errors_and_reset(renderer, teapot_path):
    sc = scenes.decal_scene(teapot_path, 48, 48)
    with_geom, tl = image(renderer, sc)
    handles = renderer.mesh_handles
    # a mesh of another vertex count
    p = np.zeros((3, 3), np.float32)
    p[1, 0] = p[2, 1] = 1.0
    small = renderer.mesh(p, np.array([[0, 1, 2]], np.uint32))
    with pytest.raises(AphError) as e:
        renderer.set_layer_geometry(tl, [[small]], 1)
    assert e.value.code == capi.APH_E_INVALID
    with pytest.raises(AphError) as e:
        renderer.set_layer_geometry(tl, [[1 << 40]], 1)
    assert e.value.code == capi.APH_E_INVALID
    # setting the layers again clears their geometry
    renderer.set_layers(tl, sc.layers, 1)
    tgt = renderer.target(48, 48)
    renderer.render([renderer.request(tgt, tl, sc.frame, 0, 2)], capi.APH_MODE_WAVEFRONT)
    renderer.wait()
    cleared = renderer.readback(tgt, 48, 48)
    plain = scenes.decal_scene(teapot_path, 48, 48)
    plain.layer_geometry = None
    assert np.array_equal(cleared, image(renderer, plain)[0])
    renderer.set_layer_geometry(tl, [[handles[1]]], 1)
    renderer.render([renderer.request(tgt, tl, sc.frame, 0, 2)], capi.APH_MODE_WAVEFRONT)
    renderer.wait()
    assert np.array_equal(renderer.readback(tgt, 48, 48), with_geom)
    # a TLAS without layers
    sc2 = scenes.teapot_scene(teapot_path, 16, 16)
    tl2 = renderer.load_scene(sc2)
    with pytest.raises(AphError) as e:
        renderer.set_layer_geometry(tl2, [[0]], 1)
    assert e.value.code == capi.APH_E_STATE
