"""Scenes read from disk (OBJ + MTL, Radiance HDR) rendered on the GPU against the oracle.

The OBJ translation follows Plasma's loader (Applications/Plasma/OBJLoader.cpp:77-470) and
the environment image is read as Plasma reads its .hdr lat-long files (Plasma.cpp:1341)."""
import numpy as np
import pytest

from aurora_amd import capi, scenes
from test_gpu_render import compare, render_both

pytestmark = pytest.mark.gpu


def write_obj(path, pos, nrm, idx):
    with open(path, "w") as f:
        f.writelines(f"v {x:.9g} {y:.9g} {z:.9g}\n" for x, y, z in pos)
        f.writelines(f"vn {x:.9g} {y:.9g} {z:.9g}\n" for x, y, z in nrm)
        f.writelines(f"f {a + 1}//{a + 1} {b + 1}//{b + 1} {c + 1}//{c + 1}\n" for a, b, c in idx)


def write_hdr(path, rgb):
    """Flat-scanline Radiance RGBE file of an (h, w, 3) float image."""
    h, w, _ = rgb.shape
    v = rgb.max(axis=-1)
    m, e = np.frexp(v)
    scale = np.where(v > 1e-32, m * 256.0 / np.maximum(v, 1e-38), 0.0)
    px = np.zeros((h, w, 4), np.uint8)
    px[..., :3] = np.clip(rgb * scale[..., None], 0, 255).astype(np.uint8)
    px[..., 3] = np.where(v > 1e-32, e + 128, 0).astype(np.uint8)
    with open(path, "wb") as f:
        f.write(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n" + f"-Y {h} +X {w}\n".encode() + px.tobytes())


def test_teapot_read_from_obj_matches_the_fixture(renderer, teapot_path, tmp_path):
    base = scenes.teapot_scene(teapot_path, 96, 96)
    pos, nrm, idx = base.meshes[0]
    unit = nrm / np.linalg.norm(nrm, axis=1, keepdims=True)
    write_obj(tmp_path / "teapot.obj", pos, unit, idx)
    sc = scenes.obj_scene(str(tmp_path / "teapot.obj"), width=96, height=96, spp=4)
    lp, ln, li = sc.meshes[0]
    # the same triangles over the same points (first-use vertex order)
    assert np.array_equal(lp[li], pos[idx].astype(np.float32))
    assert np.allclose(ln[li], unit[idx], atol=1e-6)
    for mode in (capi.APH_MODE_MEGAKERNEL, capi.APH_MODE_WAVEFRONT):
        gpu, cpu = render_both(renderer, sc, 4, mode=mode)
        compare(gpu, cpu)
    # and against the fixture scene under the same (default) material: a shape without
    # usemtl gets the renderer's default material, as in Plasma
    base.materials = [capi.default_material()]
    ref, _ = render_both(renderer, base, 4)
    compare(gpu, ref)


MTL = """newmtl glass
Kd 0.9 0.9 0.9
Ks 1 1 1
Ni 1.5
d 0.2
Tf 0.8 0.9 1
Pr 0.05

newmtl copper
Kd 0.95 0.64 0.54
Ks 1 1 1
Pm 1
Pr 0.3

newmtl lamp
Kd 0 0 0
Ke 1 0.8 0.6
"""


def test_obj_materials_and_hdr_environment_match_oracle(renderer, teapot_path, tmp_path):
    """Three teapots under MTL materials (transmissive, metallic, emissive) lit by an HDR sky
    with a bright spot, read through aph_obj_load and aph_hdr_load, in both render modes."""
    base = scenes.teapot_scene(teapot_path)
    pos, nrm, idx = base.meshes[0]
    (tmp_path / "m.mtl").write_text(MTL)
    with open(tmp_path / "three.obj", "w") as f:
        f.write("mtllib m.mtl\n")
        n = len(pos)
        for k, (name, dx) in enumerate((("glass", -2.2), ("copper", 0.0), ("lamp", 2.2))):
            f.write(f"o {name}\nusemtl {name}\n")
            f.writelines(f"v {x + dx:.9g} {y:.9g} {z:.9g}\n" for x, y, z in pos)
            # no normals: the loader's angle-weighted ones
            f.writelines(f"f {a + 1 + k * n} {b + 1 + k * n} {c + 1 + k * n}\n" for a, b, c in idx)
    h, w = 32, 64
    theta = (np.arange(h) + 0.5) / h * np.pi
    sky = np.zeros((h, w, 3), np.float32)
    sky[...] = (0.3 + 0.7 * np.cos(theta / 2) ** 2)[:, None, None] * np.array([0.5, 0.7, 1.0])
    sky[5:8, 10:14] = [40.0, 35.0, 30.0]
    write_hdr(tmp_path / "sky.hdr", sky)
    sc = scenes.obj_scene(str(tmp_path / "three.obj"), str(tmp_path / "sky.hdr"), width=128, height=64,
                          spp=4, eye=(0.0, 1.0, -7.0))
    assert [m.emission for m in sc.materials] == [0.0, 0.0, 1.0]
    read = capi.load_hdr(str(tmp_path / "sky.hdr"))
    # RGBE keeps 8 bits per channel under the pixel's largest channel's exponent
    assert np.all(np.abs(read - sky) <= sky.max(axis=-1, keepdims=True) / 128)
    for mode in (capi.APH_MODE_MEGAKERNEL, capi.APH_MODE_WAVEFRONT):
        gpu, cpu = render_both(renderer, sc, 4, mode=mode)
        compare(gpu, cpu, min_frac=0.98)
        assert gpu[..., 3].mean() > 0.05


def test_gltf_scene_with_camera_matches_oracle(renderer, teapot_path, tmp_path):
    """A binary glTF (aph_gltf_load, Plasma's glTFLoader.cpp translation): the teapot under
    two nodes of a hierarchy (one mesh, two instances), a ground quad without normals,
    pbrMetallicRoughness materials and a primitive without one, seen from the file's
    perspective camera under an HDR sky, in both render modes against the oracle."""
    from test_gltf import Doc

    base = scenes.teapot_scene(teapot_path)
    pos, nrm, idx = base.meshes[0]
    d = Doc()
    d.j["materials"] = [{"pbrMetallicRoughness": {"baseColorFactor": [0.9, 0.3, 0.2, 1],
                                                  "metallicFactor": 0.0, "roughnessFactor": 0.4}},
                        {"pbrMetallicRoughness": {"baseColorFactor": [0.95, 0.8, 0.4, 1],
                                                  "metallicFactor": 1.0, "roughnessFactor": 0.25}}]
    tea = d.mesh([{"attributes": {"POSITION": d.accessor(pos.astype(np.float32)),
                                  "NORMAL": d.accessor(nrm.astype(np.float32))},
                   "indices": d.accessor(idx.astype(np.uint16).reshape(-1), ctype=5123, typ="SCALAR"),
                   "material": 0}])
    tea_gold = d.mesh([{"attributes": {"POSITION": d.j["meshes"][0]["primitives"][0]["attributes"]["POSITION"],
                                       "NORMAL": d.j["meshes"][0]["primitives"][0]["attributes"]["NORMAL"]},
                        "indices": d.j["meshes"][0]["primitives"][0]["indices"], "material": 1}])
    g = np.array([[-6, 0, -6], [6, 0, -6], [6, 0, 6], [-6, 0, 6]], np.float32)
    floor = d.mesh([{"attributes": {"POSITION": d.accessor(g)},
                     "indices": d.accessor(np.array([0, 2, 1, 0, 3, 2], np.uint8), ctype=5121, typ="SCALAR")}])
    s = float(np.sqrt(0.5))
    left = d.node(root=False, mesh=tea, translation=[-1.6, 0, 0])
    right = d.node(root=False, mesh=tea_gold, translation=[1.6, 0, 0], rotation=[0, s, 0, s], scale=[1.2, 1.2, 1.2])
    d.node(children=[left, right], translation=[0, 0.0, 0.5])
    d.node(mesh=floor)
    d.j["cameras"] = [{"type": "perspective", "perspective": {"yfov": 0.7, "aspectRatio": 2.0, "znear": 0.1,
                                                                "zfar": 100}}]
    # camera at (0, 2, 8) pitched down by atan(1.5 / 8) about x
    a = -np.arctan2(1.5, 8.0) / 2
    d.node(camera=0, translation=[0, 2, 8], rotation=[float(np.sin(a)), 0, 0, float(np.cos(a))])
    d.write_glb(str(tmp_path / "scene.glb"))
    h, w = 32, 64
    theta = (np.arange(h) + 0.5) / h * np.pi
    sky = np.zeros((h, w, 3), np.float32)
    sky[...] = (0.3 + 0.7 * np.cos(theta / 2) ** 2)[:, None, None] * np.array([0.6, 0.7, 1.0])
    sky[4:7, 40:44] = [30.0, 28.0, 25.0]
    write_hdr(tmp_path / "sky.hdr", sky)
    sc = scenes.gltf_scene(str(tmp_path / "scene.glb"), str(tmp_path / "sky.hdr"), width=128, height=64, spp=4,
                           camera=0)
    assert len(sc.meshes) == 3 and len(sc.instances) == 3
    assert len(sc.materials) == 3  # two glTF materials + the default for the floor
    for mode in (capi.APH_MODE_MEGAKERNEL, capi.APH_MODE_WAVEFRONT):
        gpu, cpu = render_both(renderer, sc, 4, mode=mode)
        compare(gpu, cpu, min_frac=0.98)
        assert gpu[..., 3].mean() > 0.3  # teapots and floor in view
