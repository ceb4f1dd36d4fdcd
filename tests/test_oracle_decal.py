"""Pins the oracle's layer geometry (computeLayerUV, GlobalBufferAccessors.slang:116-127, in
the layer loop of HitShaderEntryPoints.slang:84-113) by invariants that hold bit for bit:

  * a layer reading texture coordinates U from its own geometry renders as the same layer
    on an instance whose mesh carries U itself (the base material untextured);
  * the base material reads the instance's coordinates even after a layer with geometry of
    its own was tried and passed through;
  * layer geometry carrying the instance's own coordinates changes nothing.
"""
import os

import numpy as np

import pyoracle
from aurora_amd import capi, scenes

TEAPOT = os.path.join(os.path.dirname(__file__), "golden", "teapot.npz")
W = H = 40


def render(sc, spp=2):
    return pyoracle.OracleScene(sc).render(sc.frame, W, H, 0, spp)[0]


def decal(**kw):
    sc = scenes.decal_scene(TEAPOT, W, H)
    for k, v in kw.items():
        setattr(sc, k, v)
    return sc


def untextured_base(sc):
    base = scenes._material((0.8, 0.8, 0.8), 0.0, 0.3)
    sc.materials = [base, sc.materials[1]]
    sc.textures = {k: v for k, v in sc.textures.items() if k[0] != 0}
    return sc


def test_layer_reads_its_own_geometry_coordinates():
    a = untextured_base(decal())
    proj = a.uvs[1]
    b = untextured_base(decal())
    b.meshes = [b.meshes[0]]
    b.uvs = {0: proj}
    b.layer_geometry = None
    ia, ib = render(a), render(b)
    assert np.array_equal(ia, ib)
    # and the coordinates matter: the layer on the generated coordinates differs
    c = untextured_base(decal())
    c.layer_geometry = None
    assert not np.array_equal(ia, render(c))


def test_base_reads_the_instance_coordinates_after_a_layer():
    def half_clear(sc):
        lay = scenes._material((0.2, 0.9, 0.3), 0.0, 0.2)
        lay.opacity[:] = (0.5, 0.5, 0.5)
        sc.materials = [sc.materials[0], lay]
        sc.textures = {k: v for k, v in sc.textures.items() if k[0] == 0}
        return sc

    e = half_clear(decal())               # the layer has geometry of its own
    f = half_clear(decal(layer_geometry=None))
    assert np.array_equal(render(e), render(f))


def test_layer_geometry_with_the_instance_coordinates_changes_nothing():
    g = decal()
    nv = len(g.meshes[0][0])
    g.uvs = {1: np.repeat((np.arange(nv, dtype=np.float32) / np.float32(nv))[:, None], 2, axis=1)}
    h = decal(layer_geometry=None)
    assert np.array_equal(render(g), render(h))
