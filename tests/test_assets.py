"""On-disk inputs: the OBJ + MTL translation of Plasma's loader
(Applications/Plasma/OBJLoader.cpp:77-470) and the Radiance HDR reader (Plasma.cpp:1341),
host code of libaph.so called without a GPU.

Expectations come from the file formats and closed forms, not from the loader's code:
hand-made RGBE bytes with known decoded values, a cube whose corner normals point into
their octants, quads whose shorter diagonal is known, and the
sRGB transfer function the MTL colours are linearized with (pch.h:125-128 approximates it)."""
import math
import os

import numpy as np
import pytest

from aurora_amd import capi


@pytest.fixture(scope="module", autouse=True)
def _lib():
    if not os.path.exists(capi.LIB_PATH):
        pytest.skip("libaph.so not built")
    capi.load()


def write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


CUBE = """# unit cube, no normals, quads
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
v 0 0 1
v 1 0 1
v 1 1 1
v 0 1 1
f 1 4 3 2
f 5 6 7 8
f 1 2 6 5
f 4 8 7 3
f 1 5 8 4
f 2 3 7 6
"""


def test_cube_corner_normals_point_outward(tmp_path):
    o = capi.load_obj(write(tmp_path, "cube.obj", CUBE))
    assert len(o["shapes"]) == 1 and o["materials"] == []
    sh = o["shapes"][0]
    assert sh["positions"].shape == (8, 3) and sh["indices"].shape == (12, 3) and sh["uvs"] is None
    assert sh["material"] == -1
    # Each corner's normal is a positively weighted sum of the three outward face normals
    # meeting there (whatever the angle weights), so it points into that corner's octant.
    centre = sh["positions"] - 0.5
    assert np.array_equal(np.sign(sh["normals"]), np.sign(centre))
    assert np.allclose(np.linalg.norm(sh["normals"], axis=1), 1.0, atol=1e-6)
    # Positions are the file's, each once (no normals or uvs to split them).
    assert sorted(map(tuple, sh["positions"].tolist())) == sorted(
        [(x, y, z) for x in (0.0, 1.0) for y in (0.0, 1.0) for z in (0.0, 1.0)])


def triangle_normal(p, t):
    n = np.cross(p[t[1]] - p[t[0]], p[t[2]] - p[t[0]])
    return n / np.linalg.norm(n)


def test_flat_quad_normals_equal_the_face_normal(tmp_path):
    o = capi.load_obj(write(tmp_path, "q.obj", "v 0 0 0\nv 2 0 0\nv 2 1 0\nv 0 1 0\nf 1 2 3 4\n"))
    sh = o["shapes"][0]
    for v in range(4):
        assert np.allclose(sh["normals"][v], [0, 0, 1], atol=1e-6)


def test_angle_weights_of_a_single_triangle(tmp_path):
    # One triangle: every corner normal is the face normal whatever the weights.
    o = capi.load_obj(write(tmp_path, "t.obj", "v 0 0 0\nv 3 0 0\nv 0 1 1\nf 1 2 3\n"))
    sh = o["shapes"][0]
    n = triangle_normal(sh["positions"], sh["indices"][0])
    assert np.allclose(sh["normals"], n[None, :], atol=1e-6)


def test_second_corner_weight_is_pi_minus_its_angle(tmp_path):
    """Foundation::calculateNormals (Geometry.cpp:59-77) weights the second corner of a
    face by pi minus its angle.  A vertex that is the second corner of one face and the
    first corner of another shows it: its normal is the sum of the two face normals
    weighted by (pi - a2) and a1."""
    # Triangle A = (P, Q, R) with Q second; triangle B = (Q, S, T) with Q first; A in z=0,
    # B tilted, so the two face normals differ.
    P, Q, R = (0.0, 0.0, 0.0), (2.0, 0.0, 0.0), (0.5, 1.5, 0.0)
    S, T = (3.0, 1.0, 1.0), (2.0, 1.0, -0.5)
    text = "".join(f"v {x} {y} {z}\n" for x, y, z in (P, Q, R, S, T)) + "f 1 2 3\nf 2 4 5\n"
    sh = capi.load_obj(write(tmp_path, "w.obj", text))["shapes"][0]
    p = {k: np.array(v) for k, v in zip("PQRST", (P, Q, R, S, T))}

    def angle(at, a, b):
        u, v = p[a] - p[at], p[b] - p[at]
        return math.acos(np.dot(u, v) / np.linalg.norm(u) / np.linalg.norm(v))

    nA = np.cross(p["Q"] - p["P"], p["R"] - p["P"])
    nA /= np.linalg.norm(nA)
    nB = np.cross(p["S"] - p["Q"], p["T"] - p["Q"])
    nB /= np.linalg.norm(nB)
    expect = (math.pi - angle("Q", "P", "R")) * nA + angle("Q", "S", "T") * nB
    expect /= np.linalg.norm(expect)
    q = [i for i in range(len(sh["positions"])) if np.allclose(sh["positions"][i], Q)]
    assert len(q) == 1
    assert np.allclose(sh["normals"][q[0]], expect, atol=1e-5)
    # and not the plain angle weighting
    plain = angle("Q", "P", "R") * nA + angle("Q", "S", "T") * nB
    plain /= np.linalg.norm(plain)
    assert not np.allclose(sh["normals"][q[0]], plain, atol=1e-3)


@pytest.mark.parametrize("kite,expect", [
    # 0-2 strictly shorter: (0, 1, 2), (0, 2, 3)
    ([(0, 0, 0), (2, -1, 0), (1, 0, 0), (2, 1, 0)], [(0, 1, 2), (0, 2, 3)]),
    # 1-3 shorter: (0, 1, 3), (1, 2, 3)
    ([(0, 0, 0), (1, -0.5, 0), (3, 0, 0), (1, 0.5, 0)], [(0, 1, 3), (1, 2, 3)]),
    # equal diagonals (a square): the 1-3 split
    ([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0)], [(0, 1, 3), (1, 2, 3)]),
])
def test_quads_split_along_the_shorter_diagonal(tmp_path, kite, expect):
    text = "".join(f"v {x} {y} {z}\n" for x, y, z in kite) + "f 1 2 3 4\n"
    sh = capi.load_obj(write(tmp_path, "k.obj", text))["shapes"][0]
    # map loaded vertices back to the file's order
    file_order = [tuple(map(float, k)) for k in kite]
    order = [file_order.index(tuple(p)) for p in sh["positions"].tolist()]
    tris = [tuple(order[i] for i in t) for t in sh["indices"].tolist()]
    assert tris == expect


def test_pentagon_is_fanned_and_covers_its_area(tmp_path):
    ang = [2 * math.pi * k / 5 for k in range(5)]
    pts = [(math.cos(a), math.sin(a), 0.0) for a in ang]
    text = "".join(f"v {x:.9f} {y:.9f} {z}\n" for x, y, z in pts) + "f 1 2 3 4 5\n"
    sh = capi.load_obj(write(tmp_path, "p.obj", text))["shapes"][0]
    assert sh["indices"].shape == (3, 3)
    p = sh["positions"].astype(np.float64)
    area = sum(0.5 * np.linalg.norm(np.cross(p[t[1]] - p[t[0]], p[t[2]] - p[t[0]])) for t in sh["indices"])
    assert area == pytest.approx(2.5 * math.sin(2 * math.pi / 5), rel=1e-6)


def test_vertices_are_distinct_position_normal_uv_triples(tmp_path):
    text = """v 0 0 0
v 1 0 0
v 0 1 0
v 1 1 0
vt 0.25 0.75
vt 1 0
vt 0 1
vn 0 0 2
vn 0 3 4
f 1/1/1 2/2/1 3/3/1
f 2/2/2 4/1/2 3/3/1
"""
    sh = capi.load_obj(write(tmp_path, "d.obj", text))["shapes"][0]
    # corners: (1,1,1) (2,2,1) (3,3,1) | (2,2,2) (4,1,2) (3,3,1) -> 5 distinct, in first-use order
    assert sh["positions"].shape == (5, 3)
    assert sh["indices"].tolist() == [[0, 1, 2], [3, 4, 2]]
    assert np.allclose(sh["positions"], [[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 0, 0], [1, 1, 0]])
    assert np.allclose(sh["normals"], [[0, 0, 1], [0, 0, 1], [0, 0, 1], [0, 0.6, 0.8], [0, 0.6, 0.8]])
    # V flipped: (u, 1 - v)
    assert np.allclose(sh["uvs"], [[0.25, 0.25], [1, 1], [0, 0], [1, 1], [0.25, 0.25]])


def test_negative_indices_and_slash_forms(tmp_path):
    # Negative indices count back from the last element read so far.
    text = ("v 0 0 0\nv 1 0 0\nv 0 1 0\nvn 0 0 1\nf -3//-1 -2//-1 -1//-1\n"
            "v 5 5 5\nvn 0 1 0\nf 1//1 2//-1 -1//2\n")
    sh = capi.load_obj(write(tmp_path, "n.obj", text))["shapes"]
    assert len(sh) == 1
    s = sh[0]
    assert s["indices"].tolist() == [[0, 1, 2], [0, 3, 4]]
    assert np.allclose(s["positions"], [[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 0, 0], [5, 5, 5]])
    assert np.allclose(s["normals"], [[0, 0, 1]] * 3 + [[0, 1, 0]] * 2)


def test_mixed_channels_are_rejected(tmp_path):
    text = "v 0 0 0\nv 1 0 0\nv 0 1 0\nvn 0 0 1\nf 1//1 2//1 3//1\nf 1 2 3\n"
    with pytest.raises(RuntimeError):
        capi.load_obj(write(tmp_path, "m.obj", text))


@pytest.mark.parametrize("face", ["f 1 2 4", "f 0 1 2", "f 1 2 -4", "f 1/5 2/1 3/1", "f 1//2 2//1 3//1",
                                  "f 1 2 -9223372036854775808", "f 1 2 4294967298", "f 1/-9223372036854775807 2 3"])
def test_out_of_range_indices_are_rejected(tmp_path, face):
    text = "v 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0 0\nvn 0 0 1\n" + face + "\n"
    with pytest.raises(RuntimeError):
        capi.load_obj(write(tmp_path, "bad.obj", text))


def test_missing_file_is_an_error(tmp_path):
    with pytest.raises(RuntimeError):
        capi.load_obj(str(tmp_path / "absent.obj"))
    with pytest.raises(RuntimeError):
        capi.load_hdr(str(tmp_path / "absent.hdr"))


def srgb_decode(v):
    """IEC 61966-2-1 sRGB decoding."""
    return v / 12.92 if v <= 0.04045 else ((v + 0.055) / 1.055) ** 2.4


MTL = """# two materials
newmtl red
Kd 1 0.5 0
Ks 0.2 0.2 0.2
Ni 1
d 0.25
Pr 0.3
Pm 0.7

newmtl glow
Kd 0 0 0
Ke 0.5 0.25 1
Tf 0.5 0.5 0.5
Ni 1.33
Tr 0.4
map_Kd tex/albedo.png
map_Pr -bm 1 rough.png
norm n.png

newmtl blank
map_Ke e.png
map_d /abs/alpha.png
bump b.png
"""

OBJ_WITH_MATERIALS = """mtllib scene.mtl
v 0 0 0
v 1 0 0
v 0 1 0
v 1 1 0
o first
usemtl glow
f 1 2 3
usemtl red
f 2 4 3
o second
f 1 2 4
o third
usemtl nothing_by_that_name
f 1 4 3
o fourth
usemtl red
f 1 3 4
"""


def test_material_translation(tmp_path):
    write(tmp_path, "scene.mtl", MTL)
    o = capi.load_obj(write(tmp_path, "scene.obj", OBJ_WITH_MATERIALS))
    assert o["material_names"] == ["red", "glow", "blank"]
    red, glow, blank = o["materials"]
    # sRGB colours become linear: 0 and 1 stay, 0.5 lands on the sRGB curve's 0.214
    assert list(red.base_color) == pytest.approx([1.0, srgb_decode(0.5), 0.0], abs=2e-3)
    assert red.base_color[0] == pytest.approx(1.0, abs=1e-6) and red.base_color[2] == 0.0
    assert list(red.specular_color) == pytest.approx([srgb_decode(0.2)] * 3, abs=2e-3)
    assert red.metalness == pytest.approx(0.7) and red.specular_roughness == pytest.approx(0.3)
    assert red.specular_ior == pytest.approx(1.5)          # Ni 1 (tinyobj's default) -> 1.5
    assert red.transmission == pytest.approx(0.75)          # 1 - d
    assert list(red.transmission_color) == [1.0, 1.0, 1.0]  # Tf absent (black) -> white
    assert red.emission == 0.0
    assert list(red.opacity) == [1.0, 1.0, 1.0]
    assert glow.specular_ior == pytest.approx(1.33)
    assert glow.transmission == pytest.approx(0.4)          # Tr without d
    assert list(glow.transmission_color) == pytest.approx([srgb_decode(0.5)] * 3, abs=2e-3)
    assert list(glow.emission_color) == pytest.approx([srgb_decode(0.5), srgb_decode(0.25), 1.0], abs=2e-3)
    assert glow.emission == 1.0
    assert list(glow.base_color) == [0.0, 0.0, 0.0]
    assert blank.emission == 1.0                            # an emission map alone turns it on
    assert list(blank.emission_color) == [0.0, 0.0, 0.0]
    # Untouched parameters keep Standard Surface defaults.
    default = capi.default_material()
    for name in ("base", "coat", "sheen", "subsurface", "specular", "thin_walled"):
        assert getattr(red, name) == getattr(default, name), name
    # Texture paths: relative to the MTL file's folder, options skipped, absolute kept.
    d = str(tmp_path) + "/"
    assert o["textures"][0] == {}
    assert o["textures"][1] == {"base_color": d + "tex/albedo.png", "specular_roughness": d + "rough.png",
                                "normal": d + "n.png"}
    assert o["textures"][2] == {"emission_color": d + "e.png", "opacity": "/abs/alpha.png", "normal": d + "b.png"}
    # Shapes: the first face's material; the current material carries over o / g (the
    # second shape's face follows "usemtl red"); an unknown name -> -1.
    assert [s["material"] for s in o["shapes"]] == [1, 0, -1, 0]
    assert [len(s["indices"]) for s in o["shapes"]] == [2, 1, 1, 1]


def test_d_wins_over_tr(tmp_path):
    write(tmp_path, "m.mtl", "newmtl a\nTr 0.9\nd 0.5\nnewmtl b\nd 0.5\nTr 0.9\n")
    o = capi.load_obj(write(tmp_path, "m.obj", "mtllib m.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n"))
    assert [m.transmission for m in o["materials"]] == pytest.approx([0.5, 0.5])


def test_srgb_linearization_tracks_the_srgb_curve(tmp_path):
    vals = np.linspace(0.0, 1.0, 11)
    lines = "".join(f"newmtl m{i}\nKd {v:.6f} {v:.6f} {v:.6f}\n" for i, v in enumerate(vals))
    write(tmp_path, "s.mtl", lines)
    o = capi.load_obj(write(tmp_path, "s.obj", "mtllib s.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n"))
    got = np.array([m.base_color[0] for m in o["materials"]])
    ref = np.array([srgb_decode(v) for v in vals])
    assert np.max(np.abs(got - ref)) < 5e-3
    assert np.all(np.diff(got) > 0)


# ---------------------------------------------------------------- Radiance HDR

def hdr_bytes(w, h, body, header=b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n"):
    return header + f"-Y {h} +X {w}\n".encode() + bytes(body)


def test_hdr_flat_pixels_decode_to_known_values(tmp_path):
    # RGBE (m, m, m, e) is m * 2^(e - 136): (128, 64, 32, 129) -> (1, 0.5, 0.25);
    # exponent 0 is black whatever the mantissas; (255, 1, 0, 140) -> (255 * 16, 16, 0).
    body = [128, 64, 32, 129, 9, 9, 9, 0, 255, 1, 0, 140,
            128, 128, 128, 128, 64, 0, 200, 120, 1, 2, 3, 136]
    p = tmp_path / "flat.hdr"
    p.write_bytes(hdr_bytes(3, 2, body))
    img = capi.load_hdr(str(p))
    assert img.shape == (2, 3, 3) and img.dtype == np.float32
    expect = [[[1, 0.5, 0.25], [0, 0, 0], [255 * 16, 16, 0]],
              [[0.5, 0.5, 0.5], [64 * 2.0 ** -16, 0, 200 * 2.0 ** -16], [1, 2, 3]]]
    assert img.tolist() == expect


def test_hdr_run_length_scanlines(tmp_path):
    # Width 10: new-style scanline header (2, 2, 0, 10), then each of R, G, B, E as runs
    # (count > 128: repeat the next byte count - 128 times) and literals (count <= 128).
    w = 10
    r = [128 + 10, 128]                              # R: ten 128s
    g = [3, 16, 32, 64, 128 + 7, 0]                  # G: 16 32 64, then seven 0s
    b = [10] + list(range(1, 11))                    # B: literal 1..10
    e = [128 + 4, 129, 128 + 6, 130]                 # E: four 129s, six 130s
    row0 = [2, 2, 0, w] + r + g + b + e
    # a second, flat scanline: all (128, 0, 0, 128)
    row1 = [128, 0, 0, 128] * w
    p = tmp_path / "rle.hdr"
    p.write_bytes(hdr_bytes(w, 2, row0 + row1, header=b"#?RGBE\nGAMMA=1\nFORMAT=32-bit_rle_rgbe\n\n"))
    img = capi.load_hdr(str(p))
    assert img.shape == (2, w, 3)
    gvals = [16, 32, 64] + [0] * 7
    for x in range(w):
        scale = 2.0 ** ((129 if x < 4 else 130) - 136)
        assert img[0, x].tolist() == [128 * scale, gvals[x] * scale, (x + 1) * scale]
        assert img[1, x].tolist() == [0.5, 0.0, 0.0]


@pytest.mark.parametrize("data", [
    b"P6\n2 2\n255\n",                                                        # not Radiance
    b"#?RADIANCE\nFORMAT=32-bit_rle_xyze\n\n-Y 1 +X 1\n\x80\x80\x80\x80",   # XYZ pixels
    b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n+Y 1 +X 1\n\x80\x80\x80\x80",   # flipped rows
    b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 2 +X 1\n\x80\x80\x80\x80",   # truncated
    b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 1 +X 8\n\x02\x02\x00\x09" + bytes(8),   # scanline width mismatch
    b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 1 +X 8\n\x02\x02\x00\x08\x8a\x01" + bytes(6),  # run past the end
    # 65535 x 65535 claimed by 16 bytes: shorter than the smallest coding of that image, so
    # rejected before 51 GB are allocated
    b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 65535 +X 65535\n" + b"\x02\x02\xff\xff" * 4,
])
def test_hdr_malformed_files_are_rejected(tmp_path, data):
    p = tmp_path / "bad.hdr"
    p.write_bytes(data)
    with pytest.raises(RuntimeError):
        capi.load_hdr(str(p))


def test_obj_scene_assembles_meshes_materials_and_environment(tmp_path):
    from aurora_amd import scenes
    write(tmp_path, "scene.mtl", MTL)
    obj = write(tmp_path, "scene.obj", OBJ_WITH_MATERIALS)
    p = tmp_path / "env.hdr"
    p.write_bytes(hdr_bytes(2, 1, [128, 64, 32, 129, 128, 128, 128, 130]))
    sc = scenes.obj_scene(obj, str(p), width=32, height=16, spp=2)
    assert len(sc.meshes) == 4 and len(sc.instances) == 4
    # shapes without a material share one default material appended after the file's
    assert list(sc.instances["material"]) == [1, 0, 3, 0] and len(sc.materials) == 4
    assert sc.env["light_rgb"].shape == (1, 2, 3)
    assert sc.env["light_rgb"][0, 1].tolist() == [2.0, 2.0, 2.0]
