"""aph_plasma (aurora_amd/apps/aph_plasma.cpp), the batch renderer shaped like Plasma's
command-line capture (Applications/Plasma/Plasma.cpp:1458-1519, :1585-1657), built with
g++ against include/aph.h alone.  Each run's image equals the same scene, camera and
sample range rendered from Python bit for bit (the program writes the aph_frame it
derived), and the oracle's render to the render bar; the progressive mode accumulates
the SampleCounter's frames into the same image as one call; the PNG holds the
post-processed 8-bit image."""
import os
import subprocess
import zlib

import numpy as np
import pytest

import pyoracle
from aurora_amd import capi, scenes
from test_gpu_assets import MTL, write_hdr
from test_gpu_render import compare

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "aurora_amd", "apps", "_build", "aph_plasma")


def run(*args):
    assert os.path.exists(BIN), "aurora_amd/apps/_build/aph_plasma is not built (run the build)"
    p = subprocess.run([BIN] + [str(a) for a in args], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    return p.stdout


def read_pfm(path):
    with open(path, "rb") as f:
        assert f.readline().strip() == b"PF"
        w, h = (int(x) for x in f.readline().split())
        assert float(f.readline()) < 0  # little-endian
        data = np.frombuffer(f.read(), "<f4").reshape(h, w, 3)
    return data[::-1]  # bottom row first in the file


def read_png(path):
    """RGBA8 PNG with unfiltered rows (what aph_plasma writes)."""
    raw = open(path, "rb").read()
    assert raw[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, w, h = 8, b"", None, None
    while pos < len(raw):
        n = int.from_bytes(raw[pos:pos + 4], "big")
        typ, body = raw[pos + 4:pos + 8], raw[pos + 8:pos + 8 + n]
        assert zlib.crc32(raw[pos + 4:pos + 8 + n]) == int.from_bytes(raw[pos + 8 + n:pos + 12 + n], "big")
        if typ == b"IHDR":
            w, h = int.from_bytes(body[:4], "big"), int.from_bytes(body[4:8], "big")
            assert body[8:10] == bytes([8, 6])
        elif typ == b"IDAT":
            idat += body
        pos += 12 + n
    rows = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, 4 * w + 1)
    assert (rows[:, 0] == 0).all()
    return rows[:, 1:].reshape(h, w, 4)


def read_frame(path):
    fr = capi.Frame()
    data = open(path, "rb").read()
    assert len(data) == capi.C.sizeof(capi.Frame)
    capi.C.memmove(capi.C.byref(fr), data, len(data))
    return fr


@pytest.fixture
def obj_files(teapot_path, tmp_path):
    base = scenes.teapot_scene(teapot_path)
    pos, nrm, idx = base.meshes[0]
    (tmp_path / "m.mtl").write_text(MTL)
    with open(tmp_path / "three.obj", "w") as f:
        f.write("mtllib m.mtl\n")
        n = len(pos)
        for k, (name, dx) in enumerate((("copper", -2.2), ("glass", 0.0), ("plain", 2.2))):
            f.write(f"o {name}\n" + (f"usemtl {name}\n" if name != "plain" else ""))
            f.writelines(f"v {x + dx:.9g} {y:.9g} {z:.9g}\n" for x, y, z in pos)
            f.writelines(f"f {a + 1 + k * n} {b + 1 + k * n} {c + 1 + k * n}\n" for a, b, c in idx)
    h, w = 32, 64
    theta = (np.arange(h) + 0.5) / h * np.pi
    sky = np.zeros((h, w, 3), np.float32)
    sky[...] = (0.3 + 0.7 * np.cos(theta / 2) ** 2)[:, None, None] * np.array([0.5, 0.7, 1.0])
    sky[5:8, 10:14] = [40.0, 35.0, 30.0]
    write_hdr(tmp_path / "sky.hdr", sky)
    return tmp_path / "three.obj", tmp_path / "sky.hdr"


def python_render(renderer, sc, frame, start, count, mode=capi.APH_MODE_WAVEFRONT):
    tlas = renderer.load_scene(sc)
    tgt = renderer.target(sc.width, sc.height)
    renderer.render([renderer.request(tgt, tlas, frame, start, count)], mode)
    renderer.wait()
    return renderer, tgt


def test_batch_render_of_an_obj_scene(renderer, obj_files, tmp_path):
    obj, hdr = obj_files
    w, h, spp = 96, 64, 8
    out = run(obj, "--env", hdr, "--dims", f"{w},{h}", "--output_spp", spp, "--output", tmp_path / "o.pfm",
              "--frame_out", tmp_path / "fr.bin", "--performance_output", tmp_path / "perf.txt")
    assert "Rendering completed in" in out and float(open(tmp_path / "perf.txt").read()) > 0
    img = read_pfm(tmp_path / "o.pfm")
    fr = read_frame(tmp_path / "fr.bin")
    # the fitted camera looks along (0, -0.5, -1) at the bounds' centre; the light is Plasma's
    back = np.array(fr.inv_view[8:11])
    assert np.allclose(back, np.array([0, 0.5, 1]) / np.sqrt(1.25), atol=1e-6)
    assert fr.light_count == 1 and fr.lights[0].color_intensity[3] == 2.0
    assert np.allclose(fr.lights[0].direction[:], -np.array([1, -0.5, 0]) / np.sqrt(1.25), atol=1e-6)
    sc = scenes.obj_scene(str(obj), str(hdr), width=w, height=h, spp=spp)
    sc.frame = fr
    r, tgt = python_render(renderer, sc, fr, 0, spp)
    ref = r.readback(tgt, w, h)
    np.testing.assert_array_equal(img, ref[..., :3])
    cpu, _ = pyoracle.OracleScene(sc).render(fr, w, h, 0, spp)
    compare(ref, cpu, min_frac=0.98)
    assert ref[..., 3].mean() > 0.1  # the teapots fill part of the view
    # the PNG is the post-processed image (no tone mapping, gamma on) stored as unorm8
    run(obj, "--env", hdr, "--dims", f"{w},{h}", "--output_spp", spp, "--output", tmp_path / "o.png")
    png = read_png(tmp_path / "o.png")
    post = r.post(tgt, w, h, tone_map=False, gamma=True)
    np.testing.assert_array_equal(png, np.round(np.clip(post, 0, 1) * 255).astype(np.uint8))
    # the .hdr output is the accumulation in RGBE (8 bits under the largest channel's exponent)
    run(obj, "--env", hdr, "--dims", f"{w},{h}", "--output_spp", spp, "--output", tmp_path / "o.hdr")
    rgbe = capi.load_hdr(str(tmp_path / "o.hdr"))
    assert np.all(np.abs(rgbe - ref[..., :3]) <= ref[..., :3].max(axis=-1, keepdims=True) / 128 + 1e-30)


def test_progressive_sample_counter_accumulates_the_same_image(renderer, obj_files, tmp_path):
    obj, hdr = obj_files
    w, h, spp = 64, 48, 40
    out = run(obj, "--env", hdr, "--dims", f"{w},{h}", "--output_spp", spp, "--progressive", 5, "--output",
              tmp_path / "p.pfm", "--frame_out", tmp_path / "fr.bin", "--mode", "megakernel")
    ranges = [tuple(int(x) for x in line.split()[-1].split("..")) for line in out.splitlines()
              if line.startswith("frame")]
    assert ranges[0][0] == 0 and ranges[-1][1] == spp
    assert all(a[1] == b[0] for a, b in zip(ranges, ranges[1:]))  # consecutive sample ranges
    assert len(ranges) >= 21  # one sample per frame while the counter primes
    fr = read_frame(tmp_path / "fr.bin")
    sc = scenes.obj_scene(str(obj), str(hdr), width=w, height=h, spp=spp)
    r, tgt = python_render(renderer, sc, fr, 0, spp, capi.APH_MODE_MEGAKERNEL)
    np.testing.assert_array_equal(read_pfm(tmp_path / "p.pfm"), r.readback(tgt, w, h)[..., :3])


def test_gltf_scene_with_its_camera_and_frame_dumps(renderer, teapot_path, tmp_path):
    from test_gltf import Doc

    base = scenes.teapot_scene(teapot_path)
    pos, nrm, idx = base.meshes[0]
    d = Doc()
    d.j["materials"] = [{"pbrMetallicRoughness": {"baseColorFactor": [0.2, 0.6, 0.9, 1], "metallicFactor": 0.0,
                                                  "roughnessFactor": 0.3}}]
    m = d.mesh([{"attributes": {"POSITION": d.accessor(pos.astype(np.float32)),
                                "NORMAL": d.accessor(nrm.astype(np.float32))},
                 "indices": d.accessor(idx.astype(np.uint32).reshape(-1), ctype=5125, typ="SCALAR"),
                 "material": 0}])
    d.node(mesh=m, rotation=[0, 0.38268343, 0, 0.92387953])
    d.j["cameras"] = [{"type": "perspective", "perspective": {"yfov": 0.6, "znear": 0.1, "zfar": 50}}]
    d.node(camera=0, translation=[0, 0.8, 6])
    glb = d.write_glb(str(tmp_path / "t.glb"))
    w, h, spp = 80, 60, 6
    run(glb, "--camera", glb, "--camera_id", 0, "--dims", f"{w},{h}", "--output_spp", spp, "--output",
        tmp_path / "g.pfm", "--frame_out", tmp_path / "fr.bin", "--lightIntensity", 0)
    fr = read_frame(tmp_path / "fr.bin")
    assert fr.lights[0].color_intensity[3] == 0.0
    sc = scenes.gltf_scene(glb, width=w, height=h, spp=spp, camera=0)
    np.testing.assert_allclose(np.array(fr.inv_view[:]), np.array(sc.frame.inv_view[:]), atol=1e-6)
    np.testing.assert_allclose(np.array(fr.view_size[:]), np.array(sc.frame.view_size[:]), rtol=1e-6)
    r, tgt = python_render(renderer, sc, fr, 0, spp)
    ref = r.readback(tgt, w, h)
    np.testing.assert_array_equal(read_pfm(tmp_path / "g.pfm"), ref[..., :3])
    cpu, _ = pyoracle.OracleScene(sc).render(fr, w, h, 0, spp)
    compare(ref, cpu, min_frac=0.98)
    assert ref[..., 3].mean() > 0.05
    # frame-sequence capture: render(0, 1) per frame, orbiting, one PNG per frame
    dump = tmp_path / "frames"
    dump.mkdir()
    run(glb, "--dims", f"{w},{h}", "--output_frames", 3, "--dump_frames_dir", dump, "--rotate_deg_per_frame", 20,
        "--output", tmp_path / "last.png")
    frames = [read_png(dump / f"frame_{k:03d}.png") for k in range(3)]
    assert not np.array_equal(frames[0], frames[1])  # the camera moved
    np.testing.assert_array_equal(frames[2], read_png(tmp_path / "last.png"))
