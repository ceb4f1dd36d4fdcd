"""The traversal's box test only culls conservatively (the property that keeps
closest hits independent of the BVH), and never accepts the inverted box of an
unused four-child slot (whose kEmptyChild reference is not a node index).

The device test (trace.h: makeBoxRay / slab) computes a plane's distance as
(p - o) * i in float32 with i = v_rcp_f32(d) (1 ulp), selects the near and far
plane by the sign of i, and accepts when max(near, tmin) <= min(far, tmax) times
kRobustFar.  Here it is restated in float32 (numpy) and checked against the
exact test in rational arithmetic, (p - o) / d with the exact reciprocal, on
random, grazing, axis-parallel and extreme rays, with the reciprocal perturbed
by +-1 ulp.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

F = np.float32
ROBUST_FAR = F(1.0000008)


def fmaxf(a, b):
    return b if math.isnan(a) else a if math.isnan(b) else max(a, b)


def fminf(a, b):
    return b if math.isnan(a) else a if math.isnan(b) else min(a, b)


def float_slab(lo, hi, o_lo, o_hi, inv, tmin, tmax):
    near, far = [], []
    for k in range(3):
        pos = inv[k] >= 0
        with np.errstate(all="ignore"):
            tl = F(F(lo[k] - o_lo[k]) * inv[k])
            th = F(F(hi[k] - o_hi[k]) * inv[k])
        near.append(tl if pos else th)
        far.append(th if pos else tl)
    tn = fmaxf(fmaxf(near[0], near[1]), fmaxf(near[2], F(tmin)))
    tf = fminf(fminf(far[0], far[1]), fminf(far[2], F(tmax)))
    with np.errstate(all="ignore"):
        return bool(tn <= F(tf * ROBUST_FAR))


def exact_slab(lo, hi, o, d, tmin, tmax):
    near, far = Fraction(tmin), Fraction(tmax) if math.isfinite(tmax) else None
    for k in range(3):
        lo_k, hi_k, o_k = Fraction(float(lo[k])), Fraction(float(hi[k])), Fraction(float(o[k]))
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


def check(lo, hi, o, d, tmin=0.0, tmax=math.inf):
    lo, hi, o, d = (np.asarray(v, dtype=np.float32) for v in (lo, hi, o, d))
    if not exact_slab(lo, hi, o, d, tmin, tmax):
        return 0
    with np.errstate(divide="ignore"):
        exact_inv = [F(1.0) / d[k] for k in range(3)]
    for k in range(3):
        for v in ulp_neighbours(exact_inv[k]):
            inv = list(exact_inv)
            if d[k] != 0:
                inv[k] = v
            assert float_slab(lo, hi, o, o, inv, tmin, tmax), (lo, hi, o, d, tmin, tmax, inv)
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
    # The unused slot's box (3e38, -3e38): every finite origin, every direction
    # (zero, signed zero, subnormal and huge components included) rejects it.
    rng = np.random.default_rng(19)
    e = F(3.0e38)
    lo, hi = np.full(3, e, np.float32), np.full(3, -e, np.float32)
    for _ in range(3000):
        o = rand_vec(rng, 10.0 ** rng.uniform(-3, 30))
        d = rand_vec(rng, 10.0 ** rng.uniform(-40, 3))
        d[rng.random(3) < 0.4] = F(0.0)
        d[rng.random(3) < 0.2] = F(-0.0)
        with np.errstate(divide="ignore", over="ignore"):
            inv = [F(1.0) / d[k] for k in range(3)]
        assert not float_slab(lo, hi, o, o, inv, 0.0, math.inf), (o, d)
        assert not float_slab(lo, hi, o, o, inv, 0.01, F(1e30)), (o, d)
