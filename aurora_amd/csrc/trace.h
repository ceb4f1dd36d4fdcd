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
    double t, u, v;
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

struct BoxRay {
    float ox, oy, oz, ix, iy, iz;
    bool sx, sy, sz;  // direction sign >= 0 per axis (selects the near plane)
};

APH_HD BoxRay makeBoxRay(f3 o, f3 d)
{
    BoxRay r;
    r.ox = o.x; r.oy = o.y; r.oz = o.z;
    r.ix = 1.0f / d.x; r.iy = 1.0f / d.y; r.iz = 1.0f / d.z;  // +-0 -> +-inf
    r.sx = r.ix >= 0.0f; r.sy = r.iy >= 0.0f; r.sz = r.iz >= 0.0f;
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
    float nx = r.sx ? tlx : thx, fx = r.sx ? thx : tlx;
    float ny = r.sy ? tly : thy, fy = r.sy ? thy : tly;
    float nz = r.sz ? tlz : thz, fz = r.sz ? thz : tlz;
    float tn = fmaxf(fmaxf(nx, ny), fmaxf(nz, tmin));
    float tf = fminf(fminf(fx, fy), fminf(fz, tmax));
    tnear = tn;
    return tn <= tf * kRobustFar;
}

// Tight-box variant (no pad) for the BLAS levels.
__device__ __forceinline__ bool slabTight(const BoxRay& r, float4 lo, float4 hi, float tmin, float tmax,
                                          float& tnear)
{
    float tlx = (lo.x - r.ox) * r.ix, thx = (hi.x - r.ox) * r.ix;
    float tly = (lo.y - r.oy) * r.iy, thy = (hi.y - r.oy) * r.iy;
    float tlz = (lo.z - r.oz) * r.iz, thz = (hi.z - r.oz) * r.iz;
    float nx = r.sx ? tlx : thx, fx = r.sx ? thx : tlx;
    float ny = r.sy ? tly : thy, fy = r.sy ? thy : tly;
    float nz = r.sz ? tlz : thz, fz = r.sz ? thz : tlz;
    float tn = fmaxf(fmaxf(nx, ny), fmaxf(nz, tmin));
    float tf = fminf(fminf(fx, fy), fminf(fz, tmax));
    tnear = tn;
    return tn <= tf * kRobustFar;
}

// Watertight ray in double (Woop, Benthin, Wald 2013), built from a float ray.
struct WRay {
    double ox, oy, oz;   // origin permuted to (kx, ky, kz)
    double Sx, Sy, Sz;
    int kx, ky, kz;
};

__device__ __forceinline__ void objectRay(const float* w2o, f3 o, f3 d, f3& oo, f3& od)
{
#pragma clang fp contract(off)
    double ox = o.x, oy = o.y, oz = o.z, dx = d.x, dy = d.y, dz = d.z;
    double r[6];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        double m0 = w2o[4 * i + 0], m1 = w2o[4 * i + 1], m2 = w2o[4 * i + 2], m3 = w2o[4 * i + 3];
        r[i] = (((m0 * ox) + (m1 * oy)) + (m2 * oz)) + m3;
        r[3 + i] = ((m0 * dx) + (m1 * dy)) + (m2 * dz);
    }
    oo = f3{(float)r[0], (float)r[1], (float)r[2]};
    od = f3{(float)r[3], (float)r[4], (float)r[5]};
}

APH_HD float comp(f3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }
APH_HD float comp4(float4 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

__device__ __forceinline__ WRay setupWRay(f3 o, f3 d)
{
#pragma clang fp contract(off)
    WRay w;
    double ax = fabs((double)d.x), ay = fabs((double)d.y), az = fabs((double)d.z);
    int kz;
    if (ax >= ay && ax >= az) kz = 0;
    else if (ay >= az) kz = 1;
    else kz = 2;
    int kx = kz + 1; if (kx == 3) kx = 0;
    int ky = kx + 1; if (ky == 3) ky = 0;
    double dkz = (double)comp(d, kz);
    if (dkz < 0.0) { int t = kx; kx = ky; ky = t; }
    w.kx = kx; w.ky = ky; w.kz = kz;
    w.Sx = (double)comp(d, kx) / dkz;
    w.Sy = (double)comp(d, ky) / dkz;
    w.Sz = 1.0 / dkz;
    w.ox = (double)comp(o, kx);
    w.oy = (double)comp(o, ky);
    w.oz = (double)comp(o, kz);
    return w;
}

__device__ __forceinline__ bool triWatertight(const WRay& w, const Tri& tr, double tmin, double tmax, double& t,
                                              double& u, double& v)
{
#pragma clang fp contract(off)
    double Akx = (double)comp4(tr.v0, w.kx) - w.ox, Aky = (double)comp4(tr.v0, w.ky) - w.oy,
           Akz = (double)comp4(tr.v0, w.kz) - w.oz;
    double Bkx = (double)comp4(tr.v1, w.kx) - w.ox, Bky = (double)comp4(tr.v1, w.ky) - w.oy,
           Bkz = (double)comp4(tr.v1, w.kz) - w.oz;
    double Ckx = (double)comp4(tr.v2, w.kx) - w.ox, Cky = (double)comp4(tr.v2, w.ky) - w.oy,
           Ckz = (double)comp4(tr.v2, w.kz) - w.oz;
    double Ax = Akx - (w.Sx * Akz), Ay = Aky - (w.Sy * Akz);
    double Bx = Bkx - (w.Sx * Bkz), By = Bky - (w.Sy * Bkz);
    double Cx = Ckx - (w.Sx * Ckz), Cy = Cky - (w.Sy * Ckz);
    double U = (Cx * By) - (Cy * Bx);
    double V = (Ax * Cy) - (Ay * Cx);
    double W = (Bx * Ay) - (By * Ax);
    if ((U < 0.0 || V < 0.0 || W < 0.0) && (U > 0.0 || V > 0.0 || W > 0.0)) return false;
    double det = (U + V) + W;
    if (det == 0.0) return false;
    double Az = w.Sz * Akz, Bz = w.Sz * Bkz, Cz = w.Sz * Ckz;
    double T = ((U * Az) + (V * Bz)) + (W * Cz);
    double tt = T / det;
    if (!(tt >= tmin && tt <= tmax)) return false;
    t = tt;
    u = V / det;
    v = W / det;
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
    best.u = best.v = 0.0;
    best.inst = best.prim = 0xFFFFFFFFu;
    bool found = false;
    if (sc.ninst == 0) return false;
    // A ray with a non-finite component misses: every triangle test rejects it, but the
    // NaN-ignoring slab test would accept every box.
    if (!(isfinite(org.x) && isfinite(org.y) && isfinite(org.z) && isfinite(dir.x) && isfinite(dir.y) &&
          isfinite(dir.z)))
        return false;
    const BoxRay wr = makeBoxRay(org, dir);
    const float wpad = sc.padScale * (fabsf(org.x) + fabsf(org.y) + fabsf(org.z) + sc.extent);
    BoxRay br = wr;
    float pad = wpad;
    const BvhNode* nodes = sc.tlas;
    const Tri* tris = nullptr;
    bool inBlas = false;
    uint32_t curInst = 0;
    WRay w;
    w.ox = w.oy = w.oz = 0.0; w.Sx = w.Sy = w.Sz = 0.0; w.kx = 0; w.ky = 1; w.kz = 2;
    const double dtmin = (double)tmin;
    stk.sp = 0;
    int32_t cur = 0;
    uint32_t steps = 0;
    while (true) {
        if (++steps > kMaxSteps) { st.overflow = 1; break; }
        if (cur >= 0) {
            const BvhNode* np = nodes + cur;
            float4 a = np->a, b = np->b, c = np->c, d = np->d;
            if (STATS) st.nodes++;
            float tlim = found ? (float)best.t * kTRoundUp : tmax;
            float t0, t1;
            bool h0, h1;
            if (inBlas) {
                h0 = slabTight(br, a, b, tmin, tlim, t0);
                h1 = slabTight(br, c, d, tmin, tlim, t1);
            } else {
                h0 = slab(br, a, b, pad, tmin, tlim, t0);
                h1 = slab(br, c, d, pad, tmin, tlim, t1);
            }
            int32_t c0 = __float_as_int(a.w), c1 = __float_as_int(b.w);
            if (h0 && h1) {
                bool swap = t1 < t0;
                if (!stk.push(swap ? c0 : c1)) { st.overflow = 1; break; }
                cur = swap ? c1 : c0;
                continue;
            }
            if (h0) { cur = c0; continue; }
            if (h1) { cur = c1; continue; }
        } else if (!inBlas) {
            // TLAS leaf: one instance.  Enter its BLAS in object space.
            uint32_t ii = leafFirst(cur);
            const InstTrav* it = sc.inst + ii;
            f3 oo, od;
            objectRay(it->w2o, org, dir, oo, od);
            w = setupWRay(oo, od);
            br = makeBoxRay(oo, od);
            nodes = it->nodes;
            tris = it->tris;
            curInst = ii;
            inBlas = true;
            if (STATS) st.insts++;
            if (!stk.push(kExitMarker)) { st.overflow = 1; break; }
            cur = 0;
            continue;
        } else {
            uint32_t first = leafFirst(cur), cnt = leafCount(cur);
            for (uint32_t i = first; i < first + cnt; i++) {
                Tri tr = tris[i];
                if (STATS) st.tris++;
                double t, u, v;
                if (!triWatertight(w, tr, dtmin, (double)tmax, t, u, v)) continue;
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
            inBlas = false;
            nodes = sc.tlas;
            br = wr;
            pad = wpad;
            if (stk.sp == 0) break;
            cur = stk.pop();
        }
    }
    return found;
}

}  // namespace aph
