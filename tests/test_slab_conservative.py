"""The traversal's box test only culls conservatively (the property that keeps
closest hits independent of the BVH), and never accepts the inverted box of an
unused four-child slot for a ray with a usable direction (the traversal also
tests the slot's kEmptyChild reference, whatever the box gives).

The device test (trace.h: boxAxis / slab) computes a plane's distance as one
float32 fused multiply-add, p * i - n on a near plane and p * j - f on a far one,
with i = v_rcp_f32(d) (1 ulp), j = i * kRobustFar, n = o i + m (+ pad |i|) and
f = (o i - m (- pad |i|)) * kRobustFar, m = 8u |o i| + 2^-120; it selects the near
plane by the sign of i and accepts when max(near, tmin) <= min(far, tmax *
kRobustFar).  Here it is restated in float32 (numpy, with a correctly rounded fma
through exact rationals) and checked against the exact test in rational
arithmetic, (p - o) / d with the exact reciprocal, on random, grazing,
axis-parallel and extreme rays, with the reciprocal perturbed by +-1 ulp, and on
instance boxes widened by a pad.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

F = np.float32
ROBUST_FAR = F(1.0000008)
MARGIN = F(4.8e-7)
TINY = F(7.5e-37)
PAD_REL = F(1.00001)
F32_ROUND_UP = Fraction(2) ** 128 - Fraction(2) ** 103  # |x| from here on rounds to inf


def fmaxf(a, b):
    return b if math.isnan(a) else a if math.isnan(b) else max(a, b)


def fminf(a, b):
    return b if math.isnan(a) else a if math.isnan(b) else min(a, b)


def round_f32(r):
    """Fraction -> float32, correctly rounded (ties to even), overflowing to inf."""
    if abs(r) >= F32_ROUND_UP:
        return F(math.copysign(math.inf, r))
    with np.errstate(over="ignore"):
        f = F(float(r))
    if not np.isfinite(f):
        f = F(math.copysign(np.finfo(np.float32).max, r))
    cands = {float(f), float(np.nextafter(f, F(np.inf))), float(np.nextafter(f, F(-np.inf)))}
    cands = [F(c) for c in cands if math.isfinite(c)]

    def key(c):
        return (abs(Fraction(float(c)) - r), int(np.asarray(c, dtype=np.float32).view(np.uint32)) & 1)

    return min(cands, key=key)


def fma32(a, b, c):
    a, b, c = F(a), F(b), F(c)
    if not (np.isfinite(a) and np.isfinite(b) and np.isfinite(c)):
        with np.errstate(all="ignore"):
            return F(np.float64(a) * np.float64(b) + np.float64(c))
    return round_f32(Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c)))


def box_axis(o, i, pad):
    """boxAxis of trace.h for a given reciprocal i: (i, j, n, f)."""
    with np.errstate(all="ignore"):
        j = F(i * ROBUST_FAR)
        oi = F(F(o) * i)
        m = fma32(abs(oi), MARGIN, TINY)
        if pad > 0:
            m = F(m + F(F(F(pad) * abs(i)) * PAD_REL))
        n = F(oi + m)
        f = F(F(oi - m) * ROBUST_FAR)
    return i, j, n, f


def float_slab(lo, hi, o, inv, tmin, tmax, pad=0.0):
    near, far = [], []
    for k in range(3):
        i, j, n, f = box_axis(o[k], inv[k], pad)
        pos = i >= 0
        near.append(fma32(lo[k] if pos else hi[k], i, -n))
        far.append(fma32(hi[k] if pos else lo[k], j, -f))
    with np.errstate(all="ignore"):
        tmax_k = F(F(tmax) * ROBUST_FAR)
    tn = fmaxf(fmaxf(near[0], near[1]), fmaxf(near[2], F(tmin)))
    tf = fminf(fminf(far[0], far[1]), fminf(far[2], tmax_k))
    return bool(tn <= tf)


def exact_slab(lo, hi, o, d, tmin, tmax):
    return exact_slab_q([Fraction(float(v)) for v in lo], [Fraction(float(v)) for v in hi], o, d, tmin, tmax)


def exact_slab_q(lo, hi, o, d, tmin, tmax):
    near, far = Fraction(tmin), Fraction(tmax) if math.isfinite(tmax) else None
    for k in range(3):
        lo_k, hi_k, o_k = lo[k], hi[k], Fraction(float(o[k]))
        if d[k] == 0:
            if not (lo_k <= o_k <= hi_k):
                return False
            continue
        t1, t2 = (lo_k - o_k) / Fraction(float(d[k])), (hi_k - o_k) / Fraction(float(d[k]))
        near = max(near, min(t1, t2))
        far = max(t1, t2) if far is None else min(far, max(t1, t2))
    return far is None or near <= far


def ulp_neighbours(x):
    x = F(x)
    return [x, np.nextafter(x, F(np.inf)), np.nextafter(x, F(-np.inf))]


def check(lo, hi, o, d, tmin=0.0, tmax=math.inf, pad=0.0):
    lo, hi, o, d = (np.asarray(v, dtype=np.float32) for v in (lo, hi, o, d))
    # the exact test on the box widened by the pad (in rationals)
    if pad > 0:
        lo_x = [Fraction(float(v)) - Fraction(float(F(pad))) for v in lo]
        hi_x = [Fraction(float(v)) + Fraction(float(F(pad))) for v in hi]
        if not exact_slab_q(lo_x, hi_x, o, d, tmin, tmax):
            return 0
    elif not exact_slab(lo, hi, o, d, tmin, tmax):
        return 0
    with np.errstate(divide="ignore"):
        exact_inv = [F(1.0) / d[k] for k in range(3)]
    for k in range(3):
        for v in ulp_neighbours(exact_inv[k]):
            inv = list(exact_inv)
            if d[k] != 0:
                inv[k] = v
            assert float_slab(lo, hi, o, inv, tmin, tmax, pad), (lo, hi, o, d, tmin, tmax, inv, pad)
    return 1


def rand_vec(rng, scale):
    return (rng.standard_normal(3) * scale).astype(np.float32)


def test_random_boxes_are_never_wrongly_culled():
    rng = np.random.default_rng(7)
    hits = 0
    for _ in range(3000):
        scale = 10.0 ** rng.uniform(-3, 4)
        o = rand_vec(rng, scale)
        d = rand_vec(rng, 1.0)
        # centres near the ray (half of them) or anywhere
        c = o + F(rng.uniform(-1.0, 4.0) * scale) * d + rand_vec(rng, scale * (0.2 if rng.uniform() < 0.5 else 1.0))
        h = np.abs(rand_vec(rng, scale * rng.uniform(0.01, 1.0)))
        hits += check(c - h, c + h, o, d, tmin=float(rng.choice([0.0, 0.01])))
    assert hits > 600


def test_grazing_boxes_are_never_wrongly_culled():
    # faces through (the float neighbours of) points on the ray, boxes of zero
    # thickness, and the origin on a face
    rng = np.random.default_rng(11)
    hits = 0
    for _ in range(3000):
        scale = 10.0 ** rng.uniform(-2, 3)
        o = rand_vec(rng, scale)
        d = rand_vec(rng, 1.0)
        t = float(rng.uniform(0.0, 3.0) * scale)
        q = o.astype(np.float64) + t * d.astype(np.float64)
        qf = q.astype(np.float32)
        ext = np.abs(rand_vec(rng, scale * 0.1))
        lo, hi = qf.copy(), qf + ext
        k = int(rng.integers(0, 3))
        step = int(rng.integers(-1, 2))
        lo[k] = np.nextafter(qf[k], F(np.inf)) if step > 0 else np.nextafter(qf[k], F(-np.inf)) if step < 0 else qf[k]
        if rng.uniform() < 0.3:
            hi[k] = lo[k]  # zero thickness on one axis
        if rng.uniform() < 0.2:
            lo, o = lo.copy(), o.copy()
            o[k] = lo[k]  # origin on the face plane
        hits += check(np.minimum(lo, hi), np.maximum(lo, hi), o, d)
    assert hits > 300


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_axis_parallel_rays(axis):
    rng = np.random.default_rng(13 + axis)
    hits = 0
    for _ in range(500):
        o = rand_vec(rng, 5.0)
        d = rand_vec(rng, 1.0)
        d[axis] = F(0.0) if rng.uniform() < 0.5 else F(-0.0)
        c = o + F(rng.uniform(0.0, 8.0)) * d + rand_vec(rng, 1.0)
        h = np.abs(rand_vec(rng, 3.0))
        if rng.uniform() < 0.3:
            o[axis] = (c - h)[axis]  # on the plane
        hits += check(c - h, c + h, o, d)
    assert hits > 100


def test_subnormal_and_huge_products():
    # tiny origins times large reciprocals and the reverse: offsets near the
    # subnormal range and near overflow
    rng = np.random.default_rng(17)
    for _ in range(1000):
        o = (rng.standard_normal(3) * 10.0 ** rng.uniform(-40, -30)).astype(np.float32)
        d = (rng.standard_normal(3) * 10.0 ** rng.uniform(-3, 3)).astype(np.float32)
        c = (rng.standard_normal(3) * 10.0 ** rng.uniform(-40, -30)).astype(np.float32)
        h = np.abs(c) + np.abs(o)
        check(c - h, c + h, o, d)
    for _ in range(300):
        o = (rng.standard_normal(3) * 1e30).astype(np.float32)
        d = (rng.standard_normal(3) * 1e-6).astype(np.float32)
        c = (rng.standard_normal(3) * 1e30).astype(np.float32)
        h = np.abs(rand_vec(rng, 1e30))
        check(c - h, c + h, o, d)


def test_inverted_boxes_are_never_accepted():
    # The unused slot's box (3e38, -3e38) is rejected by the box alone for every
    # finite origin and every direction with a component whose reciprocal i is
    # finite and nonzero while o * i stays finite (zero and signed-zero components
    # elsewhere).  Otherwise every axis can be non-limiting (an overflowing o * i
    # only ever widens the test), and the traversal rejects the slot by its
    # kEmptyChild reference, which it tests in any case.
    rng = np.random.default_rng(19)
    e = F(3.0e38)
    lo, hi = np.full(3, e, np.float32), np.full(3, -e, np.float32)
    tested = 0
    for _ in range(3000):
        o = rand_vec(rng, 10.0 ** rng.uniform(-3, 30))
        d = rand_vec(rng, 10.0 ** rng.uniform(-40, 3))
        d[rng.random(3) < 0.4] = F(0.0)
        d[rng.random(3) < 0.2] = F(-0.0)
        with np.errstate(divide="ignore", over="ignore", invalid="ignore"):
            inv = [F(1.0) / d[k] for k in range(3)]
            usable = [np.isfinite(v) and v != 0 and np.isfinite(F(o[k] * v)) for k, v in enumerate(inv)]
        if not any(usable):
            continue
        tested += 1
        for pad in (0.0, 1e-3):
            assert not float_slab(lo, hi, o, inv, 0.0, math.inf, pad), (o, d, pad)
            assert not float_slab(lo, hi, o, inv, 0.01, F(1e30), pad), (o, d, pad)
    assert tested > 1000


def test_padded_instance_boxes_are_never_wrongly_culled():
    # instance boxes: the test must accept whenever the exact test accepts the box
    # widened by the pad on every side
    rng = np.random.default_rng(23)
    hits = 0
    for _ in range(2000):
        scale = 10.0 ** rng.uniform(-2, 3)
        o = rand_vec(rng, scale)
        d = rand_vec(rng, 1.0)
        c = o + F(rng.uniform(-1.0, 4.0) * scale) * d + rand_vec(rng, scale)
        h = np.abs(rand_vec(rng, scale * rng.uniform(0.01, 1.0)))
        pad = float(F(scale * 10.0 ** rng.uniform(-7, -2)))
        hits += check(c - h, c + h, o, d, tmin=float(rng.choice([0.0, 0.01])), pad=pad)
    assert hits > 150
