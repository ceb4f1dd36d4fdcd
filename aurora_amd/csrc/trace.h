// Two-level BVH traversal with an LDS short stack and an FP64 watertight
// ray/triangle test, for gfx950.
//
// Replaces the driver's TraceRay behind traceInstanceRay / traceShadowRay
// (Shaders/RayTrace.slang:185-210, 221-270).  Hit semantics, shared with the
// CPU oracle bit for bit:
//   * the world ray is mapped to each instance's object space in double
//     (w2o 3x4 from aph_inverse3x4, products summed left to right) and rounded
//     to float; every box and triangle test of that instance uses this ray;
//   * triangles: Woop-Benthin-Wald watertight test in double, no FMA
//     contraction; U,V,W of one sign (zeros allowed), det != 0,
//     tmin <= t = T/det <= tmax;
//   * closest hit = lexicographic minimum of (t, instance, primitive).
// Boxes only cull, conservatively: BLAS boxes are tight float boxes tested with
// the robust slab test of Ize (JCGT 2(2), 2013): sign-selected planes,
// NaN-ignoring min/max, far distance widened by (1 + 2*gamma3); instance boxes
// are widened by a pad that bounds the world/object rounding mismatch.
#pragma once

#include "scene.h"

namespace aph {

constexpr int kLdsStack = 16;    // stack slots per lane held in LDS
constexpr int kSpillStack = 48;  // deeper slots spill to private (scratch) memory
constexpr int kStackCap = kLdsStack + kSpillStack;
constexpr float kRobustFar = 1.0000004f;   // 1 + 2*gamma3, gamma3 = 3u/(1-3u), u = 2^-24
constexpr float kTRoundUp = 1.0000002f;    // rounds a float-converted double t up
constexpr uint32_t kMaxSteps = 1u << 22;   // traversal step cap: a corrupt BVH cannot hang a wave

struct HitRec {
    double t;
    float u, v;
    uint32_t inst, prim;
};

struct TraceStats {
    uint32_t nodes, tris, insts, overflow;
};

// Per-lane stack: the newest kLdsStack entries in LDS (slot-major, so the 64
// lanes of a wave hit 64 consecutive dwords), older entries in private memory.
struct ShortStack {
    uint32_t* lds;
    int stride;
    int sp;
    uint32_t spill[kSpillStack];

    __device__ __forceinline__ bool push(int32_t v)
    {
        if (sp >= kStackCap) return false;
        uint32_t* slot = lds + (sp & (kLdsStack - 1)) * stride;
        if (sp >= kLdsStack) spill[sp - kLdsStack] = *slot;
        *slot = (uint32_t)v;
        sp++;
        return true;
    }
    __device__ __forceinline__ int32_t pop()
    {
        sp--;
        uint32_t* slot = lds + (sp & (kLdsStack - 1)) * stride;
        int32_t v = (int32_t)*slot;
        if (sp >= kLdsStack) *slot = spill[sp - kLdsStack];
        return v;
    }
};

// Loads through the global address space: the BVH pointers come out of device
// tables, where the compiler can only prove a generic (flat) address.
typedef float aphV4f __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 gload4(const void* p)
{
    aphV4f v = *(const __attribute__((address_space(1))) aphV4f*)p;
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ Tri gloadTri(const Tri* p)
{
    const float4* q = reinterpret_cast<const float4*>(p);
    Tri t;
    t.v0 = gload4(q);
    t.v1 = gload4(q + 1);
    t.v2 = gload4(q + 2);
    return t;
}
__device__ __forceinline__ InstTrav gloadInst(const InstTrav* p)
{
    const float4* q = reinterpret_cast<const float4*>(p);
    float4 a = gload4(q), b = gload4(q + 1), c = gload4(q + 2), d = gload4(q + 3);
    InstTrav t;
    float m[12] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c.x, c.y, c.z, c.w};
    for (int i = 0; i < 12; i++) t.w2o[i] = m[i];
    t.nodes = reinterpret_cast<const BvhNode*>(((uint64_t)__float_as_uint(d.y) << 32) | __float_as_uint(d.x));
    t.tris = reinterpret_cast<const Tri*>(((uint64_t)__float_as_uint(d.w) << 32) | __float_as_uint(d.z));
    return t;
}

struct BoxRay {
    float ox, oy, oz, ix, iy, iz;  // the sign of ix/iy/iz selects the near plane
};

APH_HD BoxRay makeBoxRay(f3 o, f3 d)
{
    BoxRay r;
    r.ox = o.x; r.oy = o.y; r.oz = o.z;
    r.ix = 1.0f / d.x; r.iy = 1.0f / d.y; r.iz = 1.0f / d.z;  // +-0 -> +-inf
    return r;
}

// Robust slab test.  A zero direction component gives +-inf distances, or NaN
// when the origin lies on the plane; fmaxf/fminf ignore NaN, which treats that
// plane as not limiting - the correct answer for a ray inside the slab.
__device__ __forceinline__ bool slab(const BoxRay& r, float4 lo, float4 hi, float pad, float tmin, float tmax,
                                     float& tnear)
{
    float lx = lo.x - pad, ly = lo.y - pad, lz = lo.z - pad;
    float hx = hi.x + pad, hy = hi.y + pad, hz = hi.z + pad;
    float tlx = (lx - r.ox) * r.ix, thx = (hx - r.ox) * r.ix;
    float tly = (ly - r.oy) * r.iy, thy = (hy - r.oy) * r.iy;
    float tlz = (lz - r.oz) * r.iz, thz = (hz - r.oz) * r.iz;
    bool sx = r.ix >= 0.0f, sy = r.iy >= 0.0f, sz = r.iz >= 0.0f;
    float nx = sx ? tlx : thx, fx = sx ? thx : tlx;
    float ny = sy ? tly : thy, fy = sy ? thy : tly;
    float nz = sz ? tlz : thz, fz = sz ? thz : tlz;
    float tn = fmaxf(fmaxf(nx, ny), fmaxf(nz, tmin));
    float tf = fminf(fminf(fx, fy), fminf(fz, tmax));
    tnear = tn;
    return tn <= tf * kRobustFar;
}

// Watertight ray (Woop, Benthin, Wald 2013) built from a float ray: the
// origin stays float (the object-space origin is float-valued by definition),
// the shear constants are double.
struct WRay {
    float ox, oy, oz;    // origin permuted to (kx, ky, kz)
    double Sx, Sy, Sz;
    uint32_t perm;       // kx | ky << 2 | kz << 4
};

__device__ __forceinline__ void objectRay(const float* w2o, f3 o, f3 d, f3& oo, f3& od)
{
#pragma clang fp contract(off)
    double ox = o.x, oy = o.y, oz = o.z, dx = d.x, dy = d.y, dz = d.z;
    float r[6];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        double m0 = w2o[4 * i + 0], m1 = w2o[4 * i + 1], m2 = w2o[4 * i + 2], m3 = w2o[4 * i + 3];
        r[i] = (float)((((m0 * ox) + (m1 * oy)) + (m2 * oz)) + m3);
        r[3 + i] = (float)(((m0 * dx) + (m1 * dy)) + (m2 * dz));
    }
    oo = f3{r[0], r[1], r[2]};
    od = f3{r[3], r[4], r[5]};
}

APH_HD float comp(f3 v, uint32_t k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }
APH_HD float comp4(float4 v, uint32_t k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

__device__ __forceinline__ WRay setupWRay(f3 o, f3 d)
{
#pragma clang fp contract(off)
    WRay w;
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    uint32_t kz;
    if (ax >= ay && ax >= az) kz = 0;
    else if (ay >= az) kz = 1;
    else kz = 2;
    uint32_t kx = kz + 1; if (kx == 3) kx = 0;
    uint32_t ky = kx + 1; if (ky == 3) ky = 0;
    double dkz = (double)comp(d, kz);
    if (dkz < 0.0) { uint32_t t = kx; kx = ky; ky = t; }
    w.perm = kx | (ky << 2) | (kz << 4);
    w.Sx = (double)comp(d, kx) / dkz;
    w.Sy = (double)comp(d, ky) / dkz;
    w.Sz = 1.0 / dkz;
    w.ox = comp(o, kx);
    w.oy = comp(o, ky);
    w.oz = comp(o, kz);
    return w;
}

// Sheared 2-D vertex (x, y) and the z offset along the ray, in double.
struct Sheared {
    double x, y, z;
};
__device__ __forceinline__ Sheared shear(const WRay& w, float4 v)
{
#pragma clang fp contract(off)
    uint32_t kx = w.perm & 3u, ky = (w.perm >> 2) & 3u, kz = w.perm >> 4;
    double z = (double)comp4(v, kz) - (double)w.oz;
    Sheared s;
    s.x = ((double)comp4(v, kx) - (double)w.ox) - (w.Sx * z);
    s.y = ((double)comp4(v, ky) - (double)w.oy) - (w.Sy * z);
    s.z = z;
    return s;
}

__device__ __forceinline__ bool triWatertight(const WRay& w, const Tri& tr, double tmin, double tmax, double& t,
                                              float& u, float& v)
{
#pragma clang fp contract(off)
    Sheared A = shear(w, tr.v0), B = shear(w, tr.v1), C = shear(w, tr.v2);
    double U = (C.x * B.y) - (C.y * B.x);
    double V = (A.x * C.y) - (A.y * C.x);
    double W = (B.x * A.y) - (B.y * A.x);
    if ((U < 0.0 || V < 0.0 || W < 0.0) && (U > 0.0 || V > 0.0 || W > 0.0)) return false;
    double det = (U + V) + W;
    if (det == 0.0) return false;
    double T = ((U * (w.Sz * A.z)) + (V * (w.Sz * B.z))) + (W * (w.Sz * C.z));
    double tt = T / det;
    if (!(tt >= tmin && tt <= tmax)) return false;
    t = tt;
    u = (float)(V / det);
    v = (float)(W / det);
    return true;
}

APH_HD bool hitBetter(double t, uint32_t inst, uint32_t prim, const HitRec& b)
{
    if (t < b.t) return true;
    if (t > b.t) return false;
    if (inst != b.inst) return inst < b.inst;
    return prim < b.prim;
}

// Traces one ray.  ANY: stop at the first accepted hit (shadow rays).
template <bool ANY, bool STATS>
__device__ bool traceRay(const TraceScene& sc, f3 org, f3 dir, float tmin, float tmax, ShortStack& stk,
                         HitRec& best, TraceStats& st)
{
    best.t = (double)tmax;
    best.u = best.v = 0.0f;
    best.inst = best.prim = 0xFFFFFFFFu;
    bool found = false;
    if (sc.ninst == 0) return false;
    // A ray with a non-finite component misses: every triangle test rejects it, but the
    // NaN-ignoring slab test would accept every box.
    if (!(isfinite(org.x) && isfinite(org.y) && isfinite(org.z) && isfinite(dir.x) && isfinite(dir.y) &&
          isfinite(dir.z)))
        return false;
    const float wpad = sc.padScale * (fabsf(org.x) + fabsf(org.y) + fabsf(org.z) + sc.extent);
    BoxRay br = makeBoxRay(org, dir);
    float pad = wpad;
    const BvhNode* nodes = sc.tlas;
    const Tri* tris = nullptr;
    uint32_t curInst = 0xFFFFFFFFu;  // != 0xFFFFFFFF while inside a BLAS
    WRay w;
    w.ox = w.oy = w.oz = 0.0f;
    w.Sx = w.Sy = w.Sz = 0.0;
    w.perm = 0;
    const double dtmin = (double)tmin, dtmax = (double)tmax;
    stk.sp = 0;
    int32_t cur = 0;
    uint32_t steps = 0;
    while (true) {
        if (++steps > kMaxSteps) { st.overflow = 1; break; }
        if (cur >= 0) {
            const float4* np = reinterpret_cast<const float4*>(nodes + cur);
            float4 a = gload4(np), b = gload4(np + 1), c = gload4(np + 2), d = gload4(np + 3);
            if (STATS) st.nodes++;
            float tlim = found ? (float)best.t * kTRoundUp : tmax;
            float t0, t1;
            bool h0 = slab(br, a, b, pad, tmin, tlim, t0);
            bool h1 = slab(br, c, d, pad, tmin, tlim, t1);
            int32_t c0 = __float_as_int(a.w), c1 = __float_as_int(b.w);
            if (h0 && h1) {
                bool swap = t1 < t0;
                if (!stk.push(swap ? c0 : c1)) { st.overflow = 1; break; }
                cur = swap ? c1 : c0;
                continue;
            }
            if (h0) { cur = c0; continue; }
            if (h1) { cur = c1; continue; }
        } else if (curInst == 0xFFFFFFFFu) {
            // TLAS leaf: one instance.  Enter its BLAS in object space.
            uint32_t ii = leafFirst(cur);
            const InstTrav it = gloadInst(sc.inst + ii);
            f3 oo, od;
            objectRay(it.w2o, org, dir, oo, od);
            w = setupWRay(oo, od);
            br = makeBoxRay(oo, od);
            pad = 0.0f;
            nodes = it.nodes;
            tris = it.tris;
            curInst = ii;
            if (STATS) st.insts++;
            if (!stk.push(kExitMarker)) { st.overflow = 1; break; }
            cur = 0;
            continue;
        } else {
            uint32_t first = leafFirst(cur), cnt = leafCount(cur);
            for (uint32_t i = first; i < first + cnt; i++) {
                Tri tr = gloadTri(tris + i);
                if (STATS) st.tris++;
                double t;
                float u, v;
                if (!triWatertight(w, tr, dtmin, dtmax, t, u, v)) continue;
                uint32_t prim = __float_as_uint(tr.v0.w);
                if (!found || hitBetter(t, curInst, prim, best)) {
                    found = true;
                    best.t = t; best.u = u; best.v = v; best.inst = curInst; best.prim = prim;
                    if (ANY) return true;
                }
            }
        }
        if (stk.sp == 0) break;
        cur = stk.pop();
        if (cur == kExitMarker) {
            curInst = 0xFFFFFFFFu;
            nodes = sc.tlas;
            br = makeBoxRay(org, dir);
            pad = wpad;
            if (stk.sp == 0) break;
            cur = stk.pop();
        }
    }
    return found;
}

}  // namespace aph
