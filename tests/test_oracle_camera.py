"""Pins the oracle's camera rays (computeCameraRay, RayTrace.slang:130-182) with
closed forms: the orthographic variant (parallel rays from the view plane), the
thin lens (every lens sample of a pixel passes through the same focal point, at
focal_distance along the pinhole direction; origins uniform on the lens disc in
the right/up plane) and its RNG draw order (the lens disc takes the draw after
the pixel jitter, RayTrace.slang:136,178; sampleDisk, Sampling.slang:90-96)."""
import math

import numpy as np
import pytest

import pyoracle
from aurora_amd import scenes


def ray(fr, W, H, x, y, s):
    o = np.zeros(3, dtype=np.float32)
    d = np.zeros(3, dtype=np.float32)
    pyoracle.load().ora_camera_ray(pyoracle.C.byref(fr), W, H, x, y, s, pyoracle._p(o), pyoracle._p(d))
    return o.astype(np.float64), d.astype(np.float64)


def basis(fr):
    inv = np.array(fr.inv_view, dtype=np.float64)
    return inv[0:3], inv[4:7], -inv[8:11], inv[12:15]


def _stream(s, pix, n):
    out = np.zeros(2 * n, dtype=np.float32)
    pyoracle.load().ora_random_stream(pix, s, n, pyoracle._p(out))
    return out.reshape(n, 2).astype(np.float64)


def test_orthographic_rays_are_parallel_from_the_view_plane():
    W, H = 32, 24
    fr = scenes.look_at_frame((1.0, 2.0, -6.0), (0.0, 0.5, 0.0), (0, 1, 0), 45.0, W / H, 3)
    fr.is_ortho = 1
    fr.view_size[0], fr.view_size[1] = 8.0, 6.0
    fr.temporal_jitter = 1
    fr.jitter[0], fr.jitter[1] = 0.25, -0.125
    right, up, front, eye = basis(fr)
    for x, y in ((0, 0), (31, 23), (7, 19), (16, 12)):
        o, d = ray(fr, W, H, x, y, 3)
        nx = (x + 0.5 + 0.25) / W * 2 - 1
        ny = -((y + 0.5 - 0.125) / H * 2 - 1)
        np.testing.assert_allclose(d, front, atol=1e-7)
        np.testing.assert_allclose(o, eye + right * 4.0 * nx + up * 3.0 * ny, atol=2e-6)


def test_thin_lens_rays_meet_at_the_focal_point():
    W, H = 40, 30
    fd, R = 7.5, 0.4
    fr = scenes.look_at_frame((0.5, 1.5, -5.0), (0.0, 0.0, 2.0), (0, 1, 0), 50.0, W / H, 3)
    fr.lens_radius, fr.focal_distance = R, fd
    fr.temporal_jitter = 1  # a fixed sub-pixel position, so only the lens sample varies
    right, up, front, eye = basis(fr)
    for x, y in ((3, 4), (20, 15), (39, 0)):
        ndc = np.array([(x + 0.5) / W * 2 - 1, -((y + 0.5) / H * 2 - 1)])
        pin = front + right * fr.view_size[0] / 2 * ndc[0] + up * fr.view_size[1] / 2 * ndc[1]
        focal = eye + fd * pin / np.linalg.norm(pin)
        r2 = []
        for s in range(400):
            o, d = ray(fr, W, H, x, y, s)
            off = o - eye
            assert abs(off @ front) < 1e-5 and np.linalg.norm(off) <= R * (1 + 1e-6)
            r2.append(off @ off)
            # distance from the focal point to the ray's line
            v = focal - o
            assert np.linalg.norm(v - (v @ d) * d) < 2e-5 * fd
            assert v @ d > 0
        assert np.mean(r2) == pytest.approx(R * R / 2, rel=0.1)  # uniform on the disc


def test_lens_takes_the_draw_after_the_pixel_jitter():
    W, H = 16, 16
    R, fd = 0.3, 4.0
    fr = scenes.look_at_frame((0.0, 0.0, -3.0), (0.0, 0.0, 0.0), (0, 1, 0), 40.0, 1.0, 3)
    fr.lens_radius, fr.focal_distance = R, fd
    right, up, front, eye = basis(fr)
    for x, y, s in ((2, 3, 0), (9, 11, 5), (15, 0, 17)):
        draws = _stream(s, y * W + x, 2)
        o, _ = ray(fr, W, H, x, y, s)
        rr, phi = R * math.sqrt(draws[1][0]), 2 * math.pi * draws[1][1]
        np.testing.assert_allclose(o, eye + right * rr * math.cos(phi) + up * rr * math.sin(phi), atol=2e-6)
        # without a lens the same pixel sample comes from the eye, jittered by the first draw
        fr.lens_radius = 0.0
        o2, d2 = ray(fr, W, H, x, y, s)
        ndc = np.array([(x + draws[0][0]) / W * 2 - 1, -((y + draws[0][1]) / H * 2 - 1)])
        pin = front + right * fr.view_size[0] / 2 * ndc[0] + up * fr.view_size[1] / 2 * ndc[1]
        np.testing.assert_allclose(o2, eye, atol=1e-6)
        np.testing.assert_allclose(d2, pin / np.linalg.norm(pin), atol=2e-6)
        fr.lens_radius = R
