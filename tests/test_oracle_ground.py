"""Pins the oracle's ground plane (the DirectX backend's Shaders/GroundPlane.slang:24-290 with
the basis of DirectX/PTGroundPlane.cpp) by closed forms:

  * the ray-plane intersection against t = (p - o).n / d.n;
  * the GGX-Smith reflection sample (BSDFCommon.slang:331-367): its estimator value times
    the cosine over the pdf is F G2 / G1 (Heitz 2018, eq. 19), and its mean over the unit
    square is the directional albedo found by quadrature of the microfacet BRDF;
  * whole pixel samples: a plane with neither effect changes nothing; a white mirror
    plane shows the background in the mirrored direction; a matte shadow under a roof
    follows the environment sample the path draws after the camera jitter; and the
    reflection coin is that draw's first component against the luminance of
    1 - reflection_opacity.
"""
import ctypes as C
import math

import numpy as np
import pytest

import pyoracle
from aurora_amd import capi, scenes
from test_oracle_camera import _stream, ray


def hit(gp, org, dirv):
    pos = np.zeros(3, dtype=np.float32)
    dist = C.c_float()
    ok = pyoracle.load().ora_ground_plane_hit(C.byref(gp), pyoracle._p(np.asarray(org, dtype=np.float32)),
                                             pyoracle._p(np.asarray(dirv, dtype=np.float32)), pyoracle._p(pos),
                                             C.byref(dist))
    return ok, pos.astype(np.float64), dist.value


def test_plane_intersection_closed_form():
    rng = np.random.default_rng(5)
    for _ in range(200):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        p = rng.uniform(-2, 2, 3)
        gp = pyoracle.ground_plane_struct(dict(position=tuple(p), normal=tuple(n)))
        o = rng.uniform(-5, 5, 3)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        o32, d32, n32, p32 = (np.asarray(v, np.float32).astype(np.float64) for v in (o, d, n, p))
        dn = d32 @ n32
        t = (p32 - o32) @ n32 / dn if dn != 0 else -1.0
        ok, pos, dist = hit(gp, o, d)
        if dn >= 0 or t < 0:
            assert not ok  # travelling along the normal, or the plane behind the ray
        else:
            assert ok
            assert dist == pytest.approx(t, rel=1e-4, abs=1e-5)
            np.testing.assert_allclose(pos, o32 + d32 * t, atol=1e-4 * (1 + abs(t)))
            assert abs((pos - p32) @ n32) < 1e-4 * (1 + abs(t))


def ggx_sample(V, f0, rough, r):
    L = np.zeros(3, np.float32)
    out = np.zeros(3, np.float32)
    pdf = C.c_float()
    pyoracle.load().ora_ggx_smith_sample(pyoracle._p(np.asarray(V, np.float32)), pyoracle._p(np.asarray(f0, np.float32)),
                                         rough, pyoracle._p(np.asarray(r, np.float32)), pyoracle._p(L), C.byref(pdf),
                                         pyoracle._p(out))
    return L.astype(np.float64), pdf.value, out.astype(np.float64)


def smith_lambda(w, a):
    # Heitz 2014 eq. 72 for GGX: (-1 + sqrt(1 + a^2 tan^2 theta)) / 2
    tan2 = (w[0] ** 2 + w[1] ** 2) / w[2] ** 2
    return (-1 + math.sqrt(1 + a * a * tan2)) / 2


def schlick(f0, c):
    return f0 + (1 - f0) * (1 - c) ** 5


@pytest.mark.parametrize("theta_deg,rough", [(20, 0.2), (55, 0.5), (75, 0.8), (40, 1.0)])
def test_ggx_estimator_is_f_g2_over_g1(theta_deg, rough):
    th = math.radians(theta_deg)
    V = np.array([math.sin(th), 0.0, math.cos(th)])
    f0 = np.array([0.9, 0.6, 0.3])
    a = max(1e-4, rough * rough)
    rng = np.random.default_rng(theta_deg)
    checked = 0
    for r in rng.uniform(0, 1, (300, 2)):
        L, pdf, f = ggx_sample(V, f0, rough, r)
        assert abs(np.linalg.norm(L) - 1) < 1e-5
        if L[2] <= 0:
            assert not f.any()  # below the surface: black
            continue
        H = (V + L) / np.linalg.norm(V + L)
        # L is V mirrored about H
        np.testing.assert_allclose(L, 2 * (V @ H) * H - V, atol=1e-4)
        g1 = 1 / (1 + smith_lambda(V, a))
        g2 = 1 / (1 + smith_lambda(V, a) + smith_lambda(L, a))
        expect = schlick(f0, V @ H) * g2 / g1
        np.testing.assert_allclose(f * L[2] / pdf, expect, rtol=2e-3, atol=1e-6)
        checked += 1
    assert checked > 120


def ggx_albedo_quadrature(V, f0, a, n=800):
    """Directional albedo of the height-correlated GGX-Smith BRDF by midpoint quadrature
    over the hemisphere (theta, phi), float64."""
    th = (np.arange(n) + 0.5) / n * (math.pi / 2)
    ph = (np.arange(2 * n) + 0.5) / (2 * n) * (2 * math.pi)
    T, P = np.meshgrid(th, ph, indexing="ij")
    L = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1)
    H = L + V
    H /= np.linalg.norm(H, axis=-1, keepdims=True)
    D = 1 / (math.pi * a * a * ((H[..., 0] ** 2 + H[..., 1] ** 2) / (a * a) + H[..., 2] ** 2) ** 2)
    lamV = smith_lambda(V, a)
    tan2L = (L[..., 0] ** 2 + L[..., 1] ** 2) / L[..., 2] ** 2
    lamL = (-1 + np.sqrt(1 + a * a * tan2L)) / 2
    G2 = 1 / (1 + lamV + lamL)
    vh = H @ V
    F = f0[None, None, :] + (1 - f0[None, None, :]) * ((1 - vh) ** 5)[..., None]
    f = F * (D * G2 / (4 * V[2] * L[..., 2]))[..., None]
    dw = np.sin(T) * (math.pi / 2 / n) * (2 * math.pi / (2 * n))
    return np.sum(f * (L[..., 2] * dw)[..., None], axis=(0, 1))


@pytest.mark.parametrize("theta_deg,rough", [(30, 0.5), (65, 0.8)])
def test_ggx_sample_mean_is_the_directional_albedo(theta_deg, rough):
    th = math.radians(theta_deg)
    V = np.array([math.sin(th), 0.0, math.cos(th)])
    f0 = np.array([0.9, 0.6, 0.3])
    m = 128
    acc = np.zeros(3)
    for i in range(m):
        for j in range(m):
            L, pdf, f = ggx_sample(V, f0, rough, ((i + 0.5) / m, (j + 0.5) / m))
            if f.any():
                acc += f * abs(L[2]) / pdf
    mc = acc / (m * m)
    ref = ggx_albedo_quadrature(V, f0, rough * rough)
    np.testing.assert_allclose(mc, ref, rtol=0.01)


def test_near_zero_roughness_is_a_mirror():
    for th in (10, 45, 80):
        t = math.radians(th)
        V = np.array([math.sin(t) * 0.6, math.sin(t) * 0.8, math.cos(t)])
        for r in ((0.1, 0.2), (0.7, 0.9), (0.5, 0.5)):
            L, pdf, f = ggx_sample(V, (1.0, 1.0, 1.0), 0.0, r)
            np.testing.assert_allclose(L, [-V[0], -V[1], V[2]], atol=2e-3)
            np.testing.assert_allclose(f * L[2] / pdf, 1.0, atol=2e-3)


# ---------------------------------------------------------------- pixel samples

W, H = 24, 18


def empty_scene(ground, roof=False, trace_depth=3, eye=(0.0, 1.0, -3.0), target=(0.0, 0.0, 2.0), quads=(), env=None):
    """A scene of the gradient environment (or `env`) and the ground plane `ground`, whose
    geometry is far away, or a large roof at y = 2, plus horizontal quads
    (y, x0, x1, z0, z1)."""
    sc = scenes.Scene()
    boxes = list(quads)
    if roof:
        boxes.append((2.0, -1.0e4, 1.0e4, -1.0e4, 1.0e4))
    if boxes:
        pos = np.array([[[x0, y, z0], [x1, y, z0], [x1, y, z1], [x0, y, z1]] for y, x0, x1, z0, z1 in boxes],
                       np.float32).reshape(-1, 3)
        idx = np.array([[4 * k, 4 * k + 1, 4 * k + 2, 4 * k, 4 * k + 2, 4 * k + 3] for k in range(len(boxes))],
                       np.uint32).reshape(-1, 3)
    else:
        pos = np.array([[1e3, 1e3, 1e3], [1e3 + 1, 1e3, 1e3], [1e3, 1e3 + 1, 1e3]], np.float32)
        idx = np.array([[0, 1, 2]], np.uint32)
    sc.meshes = [(pos, None, idx)]
    sc.materials = [capi.default_material()]
    sc.instances = scenes._instances([(0, 0, scenes.affine())])
    sc.env = dict(env or {})
    sc.ground = ground
    sc.width, sc.height, sc.spp = W, H, 1
    sc.frame = scenes.look_at_frame(eye, target, (0, 1, 0), 60.0, W / H, trace_depth)
    return sc


def sample0(sc):
    """Sample 0 of every pixel: each pixel draws its own random stream."""
    img, _ = pyoracle.OracleScene(sc).render(sc.frame, W, H, 0, 1)
    return img


def background(d, env=None):
    e, _keep = pyoracle.make_env(env or {})
    out = np.zeros(3, np.float32)
    pyoracle.load().ora_env_evaluate(C.byref(e), pyoracle._p(np.asarray(d, np.float32)), 1, 0, 0, 1, 1,
                                     pyoracle._p(out))
    return out.astype(np.float64)


def env_sample(r, env=None):
    e, _keep = pyoracle.make_env(env or {})
    L = np.zeros(3, np.float32)
    out = np.zeros(3, np.float32)
    pdf = np.zeros(1, np.float32)
    pyoracle.load().ora_env_sample(C.byref(e), pyoracle._p(np.asarray(r, np.float32)), pyoracle._p(L),
                                   pyoracle._p(pdf), pyoracle._p(out))
    return L.astype(np.float64)


def luminance32(c):
    c = np.asarray(c, np.float32)
    return np.float32(np.float32(c[0] * np.float32(0.2125)) + np.float32(c[1] * np.float32(0.7154))) + \
        np.float32(c[2] * np.float32(0.0721))


def test_plane_without_effects_changes_nothing():
    sc = scenes.ground_plane_scene(pyoracle_teapot(), "mixed", 32, 24)
    sc.ground = dict(sc.ground, reflection_opacity=0.0, shadow_opacity=0.0)
    a = pyoracle.OracleScene(sc).render(sc.frame, 32, 24, 0, 2)[0]
    sc.ground = None
    b = pyoracle.OracleScene(sc).render(sc.frame, 32, 24, 0, 2)[0]
    assert np.array_equal(a, b)


def pyoracle_teapot():
    import os
    return os.path.join(os.path.dirname(__file__), "golden", "teapot.npz")


@pytest.mark.parametrize("depth", [1, 3])
def test_white_mirror_plane_shows_the_mirrored_background(depth):
    sc = empty_scene(dict(reflection_opacity=1.0, reflection_roughness=0.0, reflection_color=(1.0, 1.0, 1.0),
                          shadow_opacity=1.0), trace_depth=depth)
    img = sample0(sc)
    down = 0
    for y in range(H):
        for x in range(W):
            o, d = ray(sc.frame, W, H, x, y, 0)
            if d[1] < -0.05:  # meets the plane through the origin facing +Y, away from grazing
                down += 1
                m = d * np.array([1, -1, 1])
                # the near-delta lobe (alpha 1e-4) keeps F G2 / G1 = 1 to fp32 cancellation
                np.testing.assert_allclose(img[y, x, :3], background(m), rtol=1e-2, atol=1e-5)
            elif d[1] < 0:
                continue  # grazing: masking and fp32 cancellation in the near-delta lobe
            else:
                np.testing.assert_allclose(img[y, x, :3], background(d), rtol=1e-6, atol=1e-7)
            assert img[y, x, 3] == 0.0  # the plane alone never makes a pixel opaque
    assert down > W * H // 4


def test_matte_shadow_follows_the_environment_draw_after_the_jitter():
    c = (0.2, 0.4, 0.6)
    # a basement under and behind the viewed part of the plane, which camera rays never reach
    # but shadow rays towards a light below the plane would: those cast no shadow
    sc = empty_scene(dict(reflection_opacity=0.0, shadow_opacity=1.0, shadow_color=c), roof=True,
                     eye=(0.0, 1.0, -0.5), target=(0.0, 0.0, 1.0), quads=[(-0.5, -1.0e4, 1.0e4, -1.0e4, -0.1)])
    img = sample0(sc)
    tint = 1.0 + (np.array(c, np.float32) - 1.0) * luminance32((1.0, 1.0, 1.0))
    shadowed = 0
    for y in range(H):
        for x in range(W):
            o, d = ray(sc.frame, W, H, x, y, 0)
            assert d[1] < 0  # every pixel looks down at the plane, under the roof
            draws = _stream(0, y * W + x, 2)
            L = env_sample(draws[1])
            bg = background(d)
            P = o + d * (-o[1] / d[1])  # on the plane
            # towards the roof, and under it: blocked
            blocked = L[1] > 0 and bool(np.all(np.abs((P + L * ((2.0 - P[1]) / L[1]))[[0, 2]]) < 1.0e4))
            if blocked:
                shadowed += 1
                np.testing.assert_allclose(img[y, x, :3], bg * tint, rtol=1e-5, atol=1e-7)
            else:  # a light below the plane casts no shadow
                np.testing.assert_allclose(img[y, x, :3], bg, rtol=1e-6, atol=1e-7)
    # uniform sphere sampling of the gradient environment: half the draws point up
    n = W * H
    assert abs(shadowed - n / 2) < 4 * math.sqrt(n / 4)


def test_reflection_coin_is_the_draw_after_the_jitter():
    ro = 0.5
    sc = empty_scene(dict(reflection_opacity=ro, reflection_roughness=0.0, reflection_color=(1.0, 1.0, 1.0),
                          shadow_opacity=0.0))
    img = sample0(sc)
    t = np.float32(1.0) - np.float32(ro)
    Pr = luminance32((t, t, t))
    reflected = 0
    for y in range(H):
        for x in range(W):
            o, d = ray(sc.frame, W, H, x, y, 0)
            if d[1] >= 0:
                continue
            draws = _stream(0, y * W + x, 2)
            if d[1] > -0.05:
                continue  # grazing (see the mirror test)
            if np.float32(draws[1][0]) >= Pr:
                reflected += 1
                expect = background(d * np.array([1, -1, 1])) * (ro / (1 - Pr))
                np.testing.assert_allclose(img[y, x, :3], expect, rtol=1e-2, atol=1e-5)
            else:
                np.testing.assert_allclose(img[y, x, :3], background(d) * (t / Pr), rtol=1e-5, atol=1e-7)
    assert 0 < reflected


def test_matte_shadow_is_cast_from_the_plane_towards_the_light():
    """A light at the zenith (the top row of a light image) and a square floating at y = 0.5
    above the viewed plane, out of the camera's view: the plane is shadowed exactly in the
    square's footprint (up to the light's 2.8 degree spread)."""
    light = np.zeros((32, 64, 3), np.float32)
    light[0, :] = 5.0
    env = dict(light_rgb=light)
    c, so = (0.1, 0.2, 0.3), 0.6
    sq = (-0.3, 0.3, -0.6, 0.1)  # x0, x1, z0, z1: not over the camera
    sc = empty_scene(dict(reflection_opacity=0.0, shadow_opacity=so, shadow_color=c), eye=(0.0, 0.25, -1.0),
                     target=(0.0, 0.0, -0.7), quads=[(0.5,) + sq], env=env)
    img = sample0(sc)
    # blocked light: alpha = luminance(1 - 0) * shadow_opacity
    tint = 1.0 + (np.array(c, np.float32) - 1.0) * (luminance32((1.0, 1.0, 1.0)) * np.float32(so))
    inside = outside = 0
    for y in range(H):
        for x in range(W):
            o, d = ray(sc.frame, W, H, x, y, 0)
            assert d[1] < 0
            P = o + d * (-o[1] / d[1])
            bg = background(d, env)
            margin = 0.05
            if sq[0] + margin < P[0] < sq[1] - margin and sq[2] + margin < P[2] < sq[3] - margin:
                inside += 1
                np.testing.assert_allclose(img[y, x, :3], bg * tint, rtol=1e-5, atol=1e-7)
            elif not (sq[0] - margin < P[0] < sq[1] + margin and sq[2] - margin < P[2] < sq[3] + margin):
                outside += 1
                np.testing.assert_allclose(img[y, x, :3], bg, rtol=1e-6, atol=1e-7)
    assert inside > 20 and outside > 20


def test_mirror_reflection_leaves_from_the_plane():
    """The same floating square seen only in a white mirror plane: pixels whose mirrored ray
    from the plane point meets the square show it, the rest the mirrored background."""
    sq = (-0.3, 0.3, -0.6, 0.1)
    sc = empty_scene(dict(reflection_opacity=1.0, reflection_roughness=0.0, reflection_color=(1.0, 1.0, 1.0)),
                     eye=(0.0, 0.25, -1.0), target=(0.0, 0.0, -0.7), quads=[(0.5,) + sq])
    img = sample0(sc)
    seen = clear = 0
    for y in range(H):
        for x in range(W):
            o, d = ray(sc.frame, W, H, x, y, 0)
            if d[1] > -0.05:
                continue
            P = o + d * (-o[1] / d[1])
            m = d * np.array([1, -1, 1])
            Q = P + m * (0.5 / m[1])  # where the mirrored ray reaches y = 0.5
            margin = 0.05
            bgm = background(m)
            if sq[0] + margin < Q[0] < sq[1] - margin and sq[2] + margin < Q[2] < sq[3] - margin:
                seen += 1
                # the underside, grey under the grey light gradient, not the blue background
                assert np.abs(img[y, x, :3] - bgm).max() > 0.2 * bgm.max()
            elif not (sq[0] - margin < Q[0] < sq[1] + margin and sq[2] - margin < Q[2] < sq[3] + margin):
                clear += 1
                np.testing.assert_allclose(img[y, x, :3], bgm, rtol=2e-2, atol=1e-5)
    assert seen > 10 and clear > 10
