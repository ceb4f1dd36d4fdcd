"""Pins the oracle's Standard Surface BSDF (StandardSurfaceBSDF.slang) with
closed forms written independently from the published lobe models:
Lambert, qualitative Oren-Nayar (1994, angle form), GGX (Walter 2007) with the
height-correlated Smith G2 (Heitz 2014 eq. 99), Schlick Fresnel, Turquin
energy compensation over the Karis curve-fit albedo, the conductor F0 = r
property of the Gulbrandsen artistic IOR mapping; the sheen LUT at grid points
(data of StandardSurfaceBSDF.slang:308-337); and, for the metal lobe, Monte
Carlo consistency between VNDF sampling (Heitz 2018) and evaluation.
"""
import ctypes as C
import math

import numpy as np
import pytest

import pyoracle
from aurora_amd import capi

FRAME = pyoracle.Frame3((0, 0, 1), (1, 0, 0), (0, 1, 0))


def mat(**kw):
    m = capi.default_material()
    for k, v in kw.items():
        if isinstance(v, (tuple, list)):
            arr = getattr(m, k)
            for i in range(3):
                arr[i] = v[i]
        else:
            setattr(m, k, v)
    return m


def evaluate(m, V, L):
    out = np.zeros(3, dtype=np.float32)
    pyoracle.load().ora_bsdf_evaluate(C.byref(m), C.byref(FRAME), pyoracle._p(pyoracle.f32(*V)),
                                      pyoracle._p(pyoracle.f32(*L)), pyoracle._p(out))
    return out.astype(np.float64)


def sample(m, V, r):
    out = np.zeros(3, dtype=np.float32)
    L = np.zeros(3, dtype=np.float32)
    pdf = C.c_float()
    lobe = C.c_int()
    pyoracle.load().ora_bsdf_sample(C.byref(m), C.byref(FRAME), pyoracle._p(pyoracle.f32(*V)),
                                    pyoracle._p(pyoracle.f32(*r)), pyoracle._p(out), pyoracle._p(L), C.byref(pdf),
                                    C.byref(lobe))
    return out.astype(np.float64), L.astype(np.float64), pdf.value, lobe.value


def eval_pdf(m, V, L, r):
    out = np.zeros(3, dtype=np.float32)
    pdf = C.c_float()
    lobe = C.c_int()
    pyoracle.load().ora_bsdf_evaluate_pdf(C.byref(m), C.byref(FRAME), pyoracle._p(pyoracle.f32(*V)),
                                          pyoracle._p(pyoracle.f32(*L)), pyoracle._p(pyoracle.f32(*r)),
                                          pyoracle._p(out), C.byref(pdf), C.byref(lobe))
    return out.astype(np.float64), pdf.value, lobe.value


def hemi_dirs(rng, n, min_z=0.05):
    d = rng.normal(size=(n, 3))
    d[:, 2] = np.abs(d[:, 2])
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return d[d[:, 2] > min_z]


# independent closed forms ---------------------------------------------------

def ggx_D(nh, a):
    return a * a / (math.pi * (nh * nh * (a * a - 1) + 1) ** 2)


def smith_lambda(cos, a):
    tan2 = (1 - cos * cos) / (cos * cos)
    return (-1 + math.sqrt(1 + a * a * tan2)) / 2


def schlick(cos, f0):
    return f0 + (1 - f0) * (1 - cos) ** 5


def karis_albedo(nv, rough, f0, f90):
    c0 = np.array([-1, -0.0275, -0.572, 0.022])
    c1 = np.array([1, 0.0425, 1.04, -0.04])
    r = rough * c0 + c1
    a004 = min(r[0] * r[0], 2 ** (-9.28 * nv)) * r[0] + r[1]
    return f0 * (-1.04 * a004 + r[2]) + f90 * (1.04 * a004 + r[3])


def test_lambert_closed_form():
    rng = np.random.default_rng(1)
    c = (0.2, 0.5, 0.9)
    m = mat(base=1.0, base_color=c, specular=0.0)
    for V in hemi_dirs(rng, 30):
        for L in hemi_dirs(rng, 5):
            np.testing.assert_allclose(evaluate(m, V, L), np.array(c) * L[2] / math.pi, rtol=2e-6)
    # opposite hemispheres: no reflection
    assert (evaluate(m, (0, 0, 1), (0, 0.6, -0.8)) == 0).all()


def test_oren_nayar_angle_form():
    rng = np.random.default_rng(2)
    rough = 0.4
    m = mat(base=1.0, base_color=(1, 1, 1), specular=0.0, diffuse_roughness=rough)
    sigma = rough * math.pi
    s2 = sigma * sigma
    A = 1 - 0.5 * s2 / (s2 + 0.33)
    B = 0.45 * s2 / (s2 + 0.09)
    for V in hemi_dirs(rng, 20):
        for L in hemi_dirs(rng, 5):
            ti, tr = math.acos(L[2]), math.acos(V[2])
            dphi = math.atan2(L[1], L[0]) - math.atan2(V[1], V[0])
            alpha, beta = max(ti, tr), min(ti, tr)
            on = A + B * max(0.0, math.cos(dphi)) * math.sin(alpha) * math.tan(beta)
            np.testing.assert_allclose(evaluate(m, V, L), L[2] / math.pi * on, rtol=1e-4)


def test_dielectric_specular_closed_form():
    rng = np.random.default_rng(3)
    for rough in (0.15, 0.4, 0.8):
        m = mat(base=0.0, specular=1.0, specular_roughness=rough, specular_ior=1.5)
        a = rough * rough
        f0 = ((1.5 - 1) / (1.5 + 1)) ** 2
        for V in hemi_dirs(rng, 25, 0.1):
            for L in hemi_dirs(rng, 4, 0.1):
                H = (V + L) / np.linalg.norm(V + L)
                F = schlick(V @ H, f0)
                G = 1 / (1 + smith_lambda(V[2], a) + smith_lambda(L[2], a))
                Ess = karis_albedo(V[2], a, 1, 1)
                comp = 1 + F * (1 - Ess) / Ess
                exp = ggx_D(H[2], a) * F * G * comp / (4 * V[2])
                np.testing.assert_allclose(evaluate(m, V, L), exp, rtol=3e-4)


def test_conductor_normal_incidence_reflectivity():
    """Gulbrandsen mapping: the conductor Fresnel at normal incidence equals the reflectivity."""
    for r in (0.1, 0.5, 0.9):
        m = mat(base=1.0, base_color=(r, r, r), metalness=1.0, specular=1.0, specular_color=(1, 1, 1),
                specular_roughness=0.3)
        a = 0.09
        Ess = karis_albedo(1.0, a, 1, 1)
        V = L = np.array([0.0, 0.0, 1.0])
        comp = 1 + r * (1 - Ess) / Ess
        exp = ggx_D(1.0, a) * r * comp * (1 / (1 + 0 + 0)) / 4
        np.testing.assert_allclose(evaluate(m, V, L), exp, rtol=2e-4)


@pytest.mark.parametrize("k,j", [(3, 0), (5, 2), (0, 6), (15, 15), (8, 9)])
def test_sheen_lut_grid_values(k, j):
    table = {(3, 0): 0.391714, (5, 2): 0.322736, (0, 6): 0.695353, (15, 15): 0.147049, (8, 9): 0.259885}
    v = pyoracle.load().ora_sheen_albedo(np.float32(k / 15), np.float32(j / 15))
    assert v == pytest.approx(table[(k, j)], rel=1e-5)


def test_metal_sampling_is_consistent_with_evaluation():
    """E[f/pdf] under VNDF sampling equals the uniform-sphere integral of the evaluated BSDF."""
    rng = np.random.default_rng(4)
    m = mat(base=1.0, base_color=(0.9, 0.7, 0.4), metalness=1.0, specular_roughness=0.5)
    V = np.array([0.3, 0.1, 0.948683])
    V /= np.linalg.norm(V)
    n = 20000
    est = np.zeros(3)
    for r in rng.uniform(0, 1, (n, 2)):
        f, L, pdf, lobe = sample(m, V, r)
        assert lobe == 3
        if pdf > 0 and (f > 0).any():
            est += f / pdf
            if rng.uniform() < 0.02:  # the pdf of the sampled direction, recomputed by evaluateMaterialAndPDF
                _, pdf2, lobe2 = eval_pdf(m, V, L, r)
                assert lobe2 == 3 and pdf2 == pytest.approx(pdf, rel=1e-3)
    est /= n
    ref = np.zeros(3)
    dirs = rng.normal(size=(4 * n, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    for L in dirs:
        if L[2] > 0:
            ref += evaluate(m, V, L)
    ref *= 4 * math.pi / len(dirs)
    np.testing.assert_allclose(est, ref, rtol=0.04)


def test_lambert_white_furnace():
    rng = np.random.default_rng(5)
    m = mat(base=1.0, base_color=(1, 1, 1), specular=0.0)
    dirs = rng.normal(size=(40000, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    tot = sum(evaluate(m, (0.2, 0.0, 0.9797959), L)[0] for L in dirs if L[2] > 0)
    assert tot * 4 * math.pi / len(dirs) == pytest.approx(1.0, rel=0.02)


def test_sheen_lut_nan_reads_row_zero():
    lib = pyoracle.load()
    assert lib.ora_sheen_albedo(np.float32(np.nan), np.float32(0.0)) == lib.ora_sheen_albedo(np.float32(0.0),
                                                                                               np.float32(0.0))
    assert lib.ora_sheen_albedo(np.float32(1.0), np.float32(np.nan)) == lib.ora_sheen_albedo(np.float32(1.0),
                                                                                               np.float32(0.0))


# ---- transmission, coat, sheen, subsurface, anisotropy, lobe selection, Q3 --------------------

TRANSPARENCY, COAT, METAL, GLOSSY, TRANSMISSION, SHEEN, DIFFUSE = 1, 2, 3, 4, 5, 6, 7
LUM = np.array([0.2125, 0.7154, 0.0721])  # Colors.slang luminance weights (Rec. 709)


def f0_of(ior):
    return ((ior - 1) / (ior + 1)) ** 2


def dirs_at(thetas, phi=0.7):
    return [np.array([math.sin(t) * math.cos(phi), math.sin(t) * math.sin(phi), math.cos(t)]) for t in thetas]


def test_transmission_refracts_by_snell_and_reflects_beyond_the_critical_angle():
    """Transmission lobe (sampleMaterial, StandardSurfaceBSDF.slang:1296-1331; computeRefraction,
    BSDFCommon.slang:122-152) against Snell's law: entering (V above) the tangential component
    shrinks by 1/ior; leaving (V below) it grows by ior, and past the critical angle
    asin(1/ior) the ray reflects about the normal with the glossy lobe ID.  A delta lobe:
    f = transmission colour (the 1/|cos| and the cosine cancel), pdf 1."""
    ior = 1.5
    col = (0.9, 0.6, 0.3)
    m = mat(base=1.0, specular=0.0, transmission=1.0, transmission_color=col, specular_ior=ior)
    crit = math.asin(1 / ior)
    for V in dirs_at(np.radians([0.5, 10, 35, 60, 85])):
        f, L, pdf, lobe = sample(m, V, (0.41, 0.77))
        st = math.sqrt(V[0] ** 2 + V[1] ** 2) / ior
        exp = np.array([-V[0] / ior, -V[1] / ior, -math.sqrt(1 - st * st)])
        np.testing.assert_allclose(L, exp, atol=2e-6)
        np.testing.assert_allclose(f, col, rtol=2e-6)
        assert pdf == 1.0 and lobe == TRANSMISSION
    for t in np.radians([5, 25, 40, 41.5, 42.5, 50, 80]):
        V = dirs_at([t])[0] * np.array([1, 1, -1])  # inside the material, looking back up
        f, L, pdf, lobe = sample(m, V, (0.41, 0.77))
        if t > crit:
            np.testing.assert_allclose(L, [-V[0], -V[1], V[2]], atol=1e-7)
            assert lobe == GLOSSY
        else:
            st = math.sin(t) * ior
            np.testing.assert_allclose(L, [-V[0] * ior, -V[1] * ior, math.sqrt(1 - st * st)], atol=3e-6)
            assert lobe == TRANSMISSION
        np.testing.assert_allclose(f, col, rtol=2e-6)
        assert pdf == 1.0


def test_thin_walled_transmission_passes_straight_through():
    m = mat(base=1.0, specular=0.0, transmission=1.0, transmission_color=(0.5, 0.7, 0.9), thin_walled=1)
    for V in dirs_at(np.radians([3, 30, 70])) + [d * np.array([1, 1, -1]) for d in dirs_at(np.radians([20, 60]))]:
        f, L, pdf, lobe = sample(m, V, (0.2, 0.9))
        np.testing.assert_allclose(L, -V, atol=1e-7)
        assert lobe == TRANSMISSION and pdf == 1.0
        np.testing.assert_allclose(f, (0.5, 0.7, 0.9), rtol=2e-6)


def ggx_reflection(V, L, a, ior, X=(1.0, 0.0, 0.0), ax=None, ay=None):
    """GGX (Walter 2007; anisotropic form of Burley 2012 when ax, ay are given, axes X and
    N x X) x Schlick x height-correlated Smith G2 at the average alpha x Turquin
    compensation over the Karis albedo fit: the MaterialX dielectric reflection lobe."""
    ax = a if ax is None else ax
    ay = a if ay is None else ay
    X = np.asarray(X, dtype=np.float64)
    Y = np.cross([0.0, 0.0, 1.0], X)
    H = (V + L) / np.linalg.norm(V + L)
    D = 1 / (math.pi * ax * ay * ((H @ X / ax) ** 2 + (H @ Y / ay) ** 2 + H[2] ** 2) ** 2)
    avg = math.sqrt(ax * ay)
    F = schlick(V @ H, f0_of(ior))
    G = 1 / (1 + smith_lambda(V[2], avg) + smith_lambda(L[2], avg))
    Ess = karis_albedo(V[2], avg, 1, 1)
    comp = 1 + F * (1 - Ess) / Ess
    return D * F * G * comp / (4 * V[2]), karis_albedo(V[2], avg, f0_of(ior), 1) * comp


def test_coat_over_a_black_base_is_a_ggx_dielectric_at_the_coat_ior():
    rng = np.random.default_rng(11)
    for rough, ior, w, c in ((0.1, 1.5, 1.0, (1, 1, 1)), (0.45, 1.8, 0.6, (0.8, 0.5, 1.0))):
        m = mat(base=0.0, specular=0.0, coat=w, coat_color=c, coat_roughness=rough, coat_ior=ior)
        for V in hemi_dirs(rng, 12, 0.15):
            for L in hemi_dirs(rng, 4, 0.15):
                spec, _ = ggx_reflection(V, L, rough * rough, ior)
                np.testing.assert_allclose(evaluate(m, V, L), w * np.array(c) * spec, rtol=3e-4)


def test_coat_layers_over_lambert_with_its_albedo_and_attenuation():
    """Layering (mx_dielectric_brdf_reflection's base term, :186-220, and the coat
    attenuation of evaluateMaterial, :698-712): f = w c spec + lerp(1, c, w) lambert (1 - w E),
    with E the coat's compensated directional albedo."""
    rng = np.random.default_rng(12)
    rho, w, c, rough = 0.5, 0.7, np.array([0.9, 0.6, 0.8]), 0.3
    m = mat(base=1.0, base_color=(rho, rho, rho), specular=0.0, coat=w, coat_color=tuple(c), coat_roughness=rough)
    for V in hemi_dirs(rng, 12, 0.15):
        for L in hemi_dirs(rng, 4, 0.15):
            spec, E = ggx_reflection(V, L, rough * rough, 1.5)
            exp = w * c * spec + (1 + (c - 1) * w) * rho * L[2] / math.pi * (1 - w * E)
            np.testing.assert_allclose(evaluate(m, V, L), exp, rtol=3e-4)


def sheen_D(cos_h, r):
    """Conty & Kulla 2017 ('Production friendly microfacet sheen BRDF'), eq. 2."""
    return (2 + 1 / r) * (1 - cos_h * cos_h) ** (0.5 / r) / (2 * math.pi)


def test_sheen_brdf_closed_form():
    """Sheen over nothing: c D(theta_h) cos_l / (4 (cos_l + cos_v - cos_l cos_v)), the
    Imageworks NDF with the smooth visibility denominator (mx_sheen_brdf_reflection, :366-410)."""
    rng = np.random.default_rng(13)
    c = np.array([0.9, 0.7, 0.5])
    for r in (0.15, 0.3, 0.8):
        m = mat(base=0.0, specular=0.0, sheen=1.0, sheen_color=tuple(c), sheen_roughness=r)
        for V in hemi_dirs(rng, 12, 0.1):
            for L in hemi_dirs(rng, 4, 0.1):
                H = (V + L) / np.linalg.norm(V + L)
                exp = c * sheen_D(H[2], r) * L[2] / (4 * (L[2] + V[2] - L[2] * V[2]))
                np.testing.assert_allclose(evaluate(m, V, L), exp, rtol=2e-4)


def test_sheen_sampling_is_cosine_with_the_ndf_as_its_pdf():
    """sampleMaterial's sheen branch (:1268-1290) draws a cosine-weighted direction but
    reports the sheen NDF at the half vector as its pdf, and its value carries the cosine
    twice (mx_sheen_brdf_specular_layer's and the final one): reference behaviour, pinned
    as such.  evaluateMaterialAndPDF reports the same pdf (:1018-1036)."""
    rng = np.random.default_rng(14)
    c = np.array([0.6, 0.8, 0.4])
    r = 0.4
    m = mat(base=0.0, specular=0.0, sheen=1.0, sheen_color=tuple(c), sheen_roughness=r)
    V = np.array([0.4, -0.2, 0.0])
    V[2] = math.sqrt(1 - V @ V)
    zs = []
    for u in rng.uniform(0, 1, (4000, 2)):
        f, L, pdf, lobe = sample(m, V, u)
        assert lobe == SHEEN
        H = (V + L) / np.linalg.norm(V + L)
        # sin^2 of the half angle cancels in fp32 near the normal: relative + absolute bar
        assert pdf == pytest.approx(sheen_D(H[2], r), rel=1e-3, abs=1e-5)
        nl = max(L[2], 1.1920929e-7)
        np.testing.assert_allclose(f, c * sheen_D(H[2], r) * nl * nl / (4 * (nl + V[2] - nl * V[2])), rtol=1e-3,
                                   atol=1e-5)
        f2, pdf2, lobe2 = eval_pdf(m, V, L, u)
        assert lobe2 == SHEEN and pdf2 == pytest.approx(pdf, rel=1e-5)
        np.testing.assert_allclose(f2, f, rtol=1e-5)
        zs.append(L[2])
    zs = np.array(zs)
    assert (zs > 0).all()
    assert zs.mean() == pytest.approx(2 / 3, abs=0.01) and (zs ** 2).mean() == pytest.approx(0.5, abs=0.01)


def test_subsurface_is_a_back_side_cosine_lobe():
    """mx_subsurface_brdf_reflection (:462-473) mixed over the diffuse lobe: light from
    behind gives s * colour * |cos| / pi, light in front (1 - s) * Lambert."""
    rng = np.random.default_rng(15)
    s, ssc, bc = 0.35, np.array([0.9, 0.5, 0.2]), np.array([0.4, 0.6, 0.8])
    m = mat(base=1.0, base_color=tuple(bc), specular=0.0, subsurface=s, subsurface_color=tuple(ssc))
    for V in hemi_dirs(rng, 10, 0.1):
        for L in hemi_dirs(rng, 4, 0.1):
            np.testing.assert_allclose(evaluate(m, V, L), (1 - s) * bc * L[2] / math.pi, rtol=2e-6)
            Lb = L * np.array([1, 1, -1])
            np.testing.assert_allclose(evaluate(m, V, Lb), s * ssc * L[2] / math.pi, rtol=2e-6)


def rotate_about_z(v, ang):
    """Rodrigues' rotation of v about +z by ang (right-handed)."""
    c, s = math.cos(ang), math.sin(ang)
    return np.array([c * v[0] - s * v[1], s * v[0] + c * v[1], v[2]])


def test_anisotropic_rotated_specular_closed_form():
    """Anisotropy (mx_roughness_anisotropy, :31-56: ax = r^2 / sqrt(1 - a), ay = r^2 sqrt(1 - a))
    about the tangent rotated by rotation * 360 degrees about the normal (mx_rotate_vector3,
    :477-502), and the same for the coat (coat_anisotropy, coat_rotation)."""
    rng = np.random.default_rng(16)
    r, an = 0.5, 0.6
    aspect = math.sqrt(1 - an)
    ax, ay = min(r * r / aspect, 1.0), r * r * aspect
    for rot in (0.0, 0.125, 0.3):
        X = rotate_about_z((1.0, 0.0, 0.0), rot * 2 * math.pi)
        ms = mat(base=0.0, specular=1.0, specular_roughness=r, specular_anisotropy=an, specular_rotation=rot)
        mc = mat(base=0.0, specular=0.0, coat=1.0, coat_roughness=r, coat_anisotropy=an, coat_rotation=rot)
        for V in hemi_dirs(rng, 10, 0.2):
            for L in hemi_dirs(rng, 4, 0.2):
                spec, _ = ggx_reflection(V, L, None, 1.5, X=X, ax=ax, ay=ay)
                np.testing.assert_allclose(evaluate(ms, V, L), spec, rtol=5e-4)
                np.testing.assert_allclose(evaluate(mc, V, L), spec, rtol=5e-4)
    # a half turn leaves the lobe unchanged (the NDF is even in the tangent)
    V, L = np.array([0.3, 0.5, 0.81240384]), np.array([-0.6, 0.2, 0.77459667])
    a = evaluate(mat(base=0.0, specular_anisotropy=an, specular_rotation=0.1), V, L)
    b = evaluate(mat(base=0.0, specular_anisotropy=an, specular_rotation=0.6), V, L)
    np.testing.assert_allclose(a, b, rtol=1e-4)


def test_coat_affect_color_and_roughness_are_the_documented_substitutions():
    """coat_affect_color raises the base and subsurface colours to 1 + coat * coat_affect_color;
    coat_affect_roughness replaces the specular roughness by lerp(r, 1, coat_affect_roughness *
    coat * coat_roughness) (the prologue of evaluateMaterial, :598-629)."""
    rng = np.random.default_rng(17)
    bc, ssc, coat = np.array([0.7, 0.4, 0.2]), np.array([0.3, 0.8, 0.5]), 0.8
    g = 1 + coat * 1.0
    common = dict(base=1.0, specular=1.0, coat=coat, coat_roughness=0.4, subsurface=0.3)
    a = mat(base_color=tuple(bc), subsurface_color=tuple(ssc), coat_affect_color=1.0, **common)
    b = mat(base_color=tuple(bc ** g), subsurface_color=tuple(ssc ** g), coat_affect_color=0.0, **common)
    r0, cr = 0.2, 0.6
    c = mat(base=1.0, specular=1.0, specular_roughness=r0, coat=0.5, coat_roughness=cr, coat_affect_roughness=1.0)
    d = mat(base=1.0, specular=1.0, specular_roughness=r0 + (1 - r0) * 0.5 * cr, coat=0.5, coat_roughness=cr)
    for V in hemi_dirs(rng, 8, 0.15):
        for L in hemi_dirs(rng, 3, 0.15):
            np.testing.assert_allclose(evaluate(a, V, L), evaluate(b, V, L), rtol=2e-5)
            Lb = L * np.array([1, 1, -1])
            np.testing.assert_allclose(evaluate(a, V, Lb), evaluate(b, V, Lb), rtol=2e-5)
            np.testing.assert_allclose(evaluate(c, V, L), evaluate(d, V, L), rtol=2e-5)
    assert not np.allclose(evaluate(c, V, L), evaluate(mat(base=1.0, specular=1.0, specular_roughness=r0,
                                                           coat=0.5, coat_roughness=cr), V, L))


LOBE_MAT = dict(base=1.0, base_color=(0.6, 0.5, 0.4), coat=0.6, coat_color=(0.9, 0.8, 0.7), coat_ior=1.6,
                metalness=0.3, specular=0.8, specular_color=(1.0, 0.9, 0.8), specular_ior=1.45, sheen=0.4,
                transmission=0.5)


def lobe_probabilities(V, with_transmission=True):
    """The lobe cascade coat -> metal -> glossy -> sheen -> transmission -> diffuse
    (:1156-1355), each taken with its probability of what is left: coat by the luminance of
    coat * colour * Schlick(N.V) at the coat IOR, glossy likewise at the specular IOR."""
    p = LOBE_MAT
    cl = float(LUM @ (np.array(p["coat_color"]) * schlick(V[2], f0_of(p["coat_ior"])) * p["coat"]))
    gl = float(LUM @ (np.array(p["specular_color"]) * schlick(V[2], f0_of(p["specular_ior"])) * p["specular"]))
    rest = 1 - cl
    out = {COAT: cl, METAL: rest * p["metalness"]}
    rest *= 1 - p["metalness"]
    out[GLOSSY] = rest * gl
    rest *= 1 - gl
    out[SHEEN] = rest * p["sheen"]
    rest *= 1 - p["sheen"]
    if with_transmission:
        out[TRANSMISSION] = rest * p["transmission"]
        rest *= 1 - p["transmission"]
    out[DIFFUSE] = rest
    return out


@pytest.mark.parametrize("vz", [0.95, 0.35])
def test_lobe_selection_frequencies_over_xi(vz):
    m = mat(**LOBE_MAT)
    V = np.array([math.sqrt(1 - vz * vz), 0.0, vz])
    L = np.array([-0.3, 0.4, 0.0])
    L[2] = math.sqrt(1 - L @ L)
    n = 20000
    xs = (np.arange(n) + 0.5) / n
    for with_t, fn in ((True, lambda x: sample(m, V, (x, 0.37))[3]), (False, lambda x: eval_pdf(m, V, L, (x, 0.37))[2])):
        ids = np.array([fn(x) for x in xs])
        for lobe, p in lobe_probabilities(V, with_t).items():
            assert (ids == lobe).mean() == pytest.approx(p, abs=3e-4), (with_t, lobe)
        assert set(np.unique(ids)) <= set(lobe_probabilities(V, with_t))


def ggx_pdf(V, L, a):
    """Density of the reflected direction under VNDF sampling: D(H) G1(V) / (4 |V.z|) (Heitz 2018)."""
    H = (V + L) / np.linalg.norm(V + L)
    return ggx_D(H[2], a) / (1 + smith_lambda(V[2], a)) / (4 * V[2])


def test_evaluate_and_pdf_returns_the_pdf_of_the_selected_lobe_only():
    """Quirk Q3 (evaluateMaterialAndPDF, :909-1076): the pdf is that of the one lobe xi.x
    selects - never the mixture - and there is no transmission branch (a transmissive
    material reports the diffuse lobe and its cosine pdf)."""
    rng = np.random.default_rng(18)
    V = np.array([0.2, 0.3, 0.0])
    V[2] = math.sqrt(1 - V @ V)
    diffuse = mat(base=1.0, specular=0.0)
    metal = mat(base=1.0, metalness=1.0, specular_roughness=0.45)
    two = mat(base=1.0, specular=1.0, specular_roughness=0.35)
    trans = mat(base=1.0, specular=0.0, transmission=1.0)
    gl = float(LUM @ (np.ones(3) * schlick(V[2], f0_of(1.5))))
    for L in hemi_dirs(rng, 20, 0.1):
        for x in (0.01, 0.3, 0.7, 0.99):
            _, pdf, lobe = eval_pdf(diffuse, V, L, (x, 0.5))
            assert lobe == DIFFUSE and pdf == pytest.approx(L[2] / math.pi, rel=1e-5)
            _, pdf, lobe = eval_pdf(trans, V, L, (x, 0.5))
            assert lobe == DIFFUSE and pdf == pytest.approx(L[2] / math.pi, rel=1e-5)
            _, pdf, lobe = eval_pdf(metal, V, L, (x, 0.5))
            assert lobe == METAL and pdf == pytest.approx(ggx_pdf(V, L, 0.45 ** 2), rel=3e-4)
            _, pdf, lobe = eval_pdf(two, V, L, (x * gl, 0.5))
            assert lobe == GLOSSY and pdf == pytest.approx(ggx_pdf(V, L, 0.35 ** 2), rel=3e-4)
            _, pdf, lobe = eval_pdf(two, V, L, (gl + x * (1 - gl), 0.5))
            assert lobe == DIFFUSE and pdf == pytest.approx(L[2] / math.pi, rel=1e-5)
    # the sampling pdf of the metal lobe ignores coat_affect_roughness (specAlpha, :947), the value does not
    ca = mat(base=1.0, metalness=1.0, specular_roughness=0.3, coat=1.0, coat_roughness=0.5,
             coat_affect_roughness=1.0, coat_ior=1.0001, coat_color=(0, 0, 0))
    for L in hemi_dirs(rng, 5, 0.2):
        _, pdf, lobe = eval_pdf(ca, V, L, (0.5, 0.5))
        assert lobe == METAL and pdf == pytest.approx(ggx_pdf(V, L, 0.09), rel=3e-4)
