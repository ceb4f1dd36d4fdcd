// Guide images (AOVs) and the temporal resolve, for gfx950.
//
// AOVs: the DirectX backend's denoising guides written by the primary hit or
// miss (HitShaderEntryPoints.slang:150-215, MainEntryPoints.slang:385-405,
// motion vectors GlobalPipelineState.slang:126-172): view-space depth, normal
// + clamped roughness, base colour + metalness, and 2.5D motion (previous
// minus current pixel position; |prev viewZ| - |cur viewZ|).  One lane per
// pixel traces the camera ray of one sample index with the path tracer's own
// traversal, so the guides belong to exactly the first segment of that sample.
//
// Temporal resolve: TemporalResolve.hlsl:156-249.  3x3 YCoCg moments give a
// one-sigma box intersected with the neighbourhood min/max (the current colour
// kept inside); history is resampled with a Catmull-Rom bicubic at the
// reprojected position, clipped toward the current colour along the line
// between them, and rejected off screen, across a depth change of more than 5%
// of the view depth, when non-finite, or on reset; the current-frame weight
// grows with on-screen speed from blendAlphaMin to blendAlphaMax.
#include <hip/hip_runtime.h>

#include "pathtrace.h"
#include "temporal.h"

namespace aph {

namespace {

constexpr int kAovBlock = 128;
static_assert(kAovBlock == kStackLanes, "the traversal stack stride is the block size");

__device__ __forceinline__ float4 mul4(const float* M, float x, float y, float z, float w)
{
#pragma clang fp contract(off)
    return make_float4(((M[0] * x + M[1] * y) + M[2] * z) + M[3] * w, ((M[4] * x + M[5] * y) + M[6] * z) + M[7] * w,
                       ((M[8] * x + M[9] * y) + M[10] * z) + M[11] * w,
                       ((M[12] * x + M[13] * y) + M[14] * z) + M[15] * w);
}

// computeViewZ, GlobalPipelineState.slang:142-149: rows of invView are right,
// up, back, eye; negative in front of the camera.
__device__ __forceinline__ float viewZ(const float* invView, f3 p)
{
#pragma clang fp contract(off)
    f3 front = f3{-invView[8], -invView[9], -invView[10]};
    f3 d = f3{p.x - invView[12], p.y - invView[13], p.z - invView[14]};
    return -((d.x * front.x + d.y * front.y) + d.z * front.z);
}

// computeScreenMotion, GlobalPipelineState.slang:126-140.
__device__ __forceinline__ f2 screenMotion(const float* vp, const float* vpPrev, float x, float y, float z, float w,
                                           uint32_t W, uint32_t H)
{
#pragma clang fp contract(off)
    float4 c = mul4(vp, x, y, z, w), p = mul4(vpPrev, x, y, z, w);
    float cw = fmaxf(fabsf(c.w), 1e-6f), pw = fmaxf(fabsf(p.w), 1e-6f);
    float cx = c.x / cw, cy = c.y / cw, px = p.x / pw, py = p.y / pw;
    float sw = (float)W, sh = (float)H;
    float csx = (cx * 0.5f + 0.5f) * sw, csy = (-cy * 0.5f + 0.5f) * sh;
    float psx = (px * 0.5f + 0.5f) * sw, psy = (-py * 0.5f + 0.5f) * sh;
    return f2{psx - csx, psy - csy};
}

__global__ __launch_bounds__(kAovBlock) void kAovs(RenderScene rs, AovParams p, AovPlanes out)
{
#pragma clang fp contract(off)
    __shared__ uint32_t sStack[kLdsStack * kAovBlock];
    uint32_t spill[kSpillStack];
    const StackMem stk{sStack + threadIdx.x, spill};
    const uint32_t n = p.W * p.H;
    const uint32_t i = blockIdx.x * kAovBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t x = i % p.W, y = i / p.W;
    Rng rng = rngInit(p.sampleIndex, p.W, x, y);
    f3 org, dir;
    cameraRay(p.fp, x, y, p.W, p.H, rng, org, dir);
    HitRec h;
    TraceStats ts = {};
    bool hit = traceRay<false, false>(rs.sc, org, dir, 0.0f, INFINITY, stk, h, ts);
    if (!hit) {
        f2 m = screenMotion(p.viewProj, p.viewProjPrev, dir.x, dir.y, dir.z, 0.0f, p.W, p.H);
        out.depth[i] = p.skyViewZ;
        out.normalRoughness[i] = make_float4(0.0f, 0.0f, 0.0f, 0.1f);
        out.baseColorMetalness[i] = make_float4(0.0f, 0.0f, 0.0f, 0.1f);
        out.motion[i] = make_float4(m.x, m.y, 0.0f, 0.0f);
        return;
    }
    const float t = (float)h.t;
    f3 pos = f3{org.x + dir.x * t, org.y + dir.y * t, org.z + dir.z * t};
    float vz = viewZ(p.fp.inv_view, pos);
    const InstShade& ins = rs.ishade[h.inst];
    Shading sh;
    computeShading(ins, h.prim, h.u, h.v, sh);
    uint32_t mi = ins.material < rs.nmat ? ins.material : 0u;
    const DevMaterial& m = rs.mats[mi];
    HitUV huv(ins, h.prim, h.u, h.v);
    f3 emissionColor;
    materialTextures(rs, mi, m, huv, sh, emissionColor);  // base colour, roughness, normal map
    f3 nrm = normalize(sh.normal);
    f2 mv = screenMotion(p.viewProj, p.viewProjPrev, pos.x, pos.y, pos.z, 1.0f, p.W, p.H);
    float prevZ = viewZ(p.invViewPrev, pos);
    out.depth[i] = vz;
    out.normalRoughness[i] = make_float4(nrm.x, nrm.y, nrm.z, fmaxf(sh.specularRoughness, 0.05f));
    out.baseColorMetalness[i] = make_float4(sh.baseColor.x, sh.baseColor.y, sh.baseColor.z, m.metalness);
    out.motion[i] = make_float4(mv.x, mv.y, vz - prevZ, 0.0f);
}

// ---- temporal resolve ----
__device__ __forceinline__ f3 toYCoCg(f3 c)
{
#pragma clang fp contract(off)
    return f3{(0.25f * c.x + 0.5f * c.y) + 0.25f * c.z, 0.5f * c.x - 0.5f * c.z, (-0.25f * c.x + 0.5f * c.y) - 0.25f * c.z};
}
__device__ __forceinline__ f3 fromYCoCg(f3 c)
{
#pragma clang fp contract(off)
    float t = c.x - c.z;
    return f3{t + c.y, c.x + c.z, t - c.y};
}

__device__ __forceinline__ f3 catmullRom(const float4* hist, float2 pos, int W, int H)
{
#pragma clang fp contract(off)
    float sx = pos.x - 0.5f, sy = pos.y - 0.5f;
    float bx = floorf(sx), by = floorf(sy);
    float fx = sx - bx, fy = sy - by;
    float wx[4], wy[4];
    {
        float f = fx, f2 = f * f, f3 = f2 * f;
        wx[0] = (-0.5f * f3 + f2) - 0.5f * f;
        wx[1] = (1.5f * f3 - 2.5f * f2) + 1.0f;
        wx[2] = (-1.5f * f3 + 2.0f * f2) + 0.5f * f;
        wx[3] = 0.5f * f3 - 0.5f * f2;
    }
    {
        float f = fy, f2 = f * f, f3 = f2 * f;
        wy[0] = (-0.5f * f3 + f2) - 0.5f * f;
        wy[1] = (1.5f * f3 - 2.5f * f2) + 1.0f;
        wy[2] = (-1.5f * f3 + 2.0f * f2) + 0.5f * f;
        wy[3] = 0.5f * f3 - 0.5f * f2;
    }
    f3 r = splat(0.0f);
    // The taps clamp to the image, so a base beyond [-4, size + 4] reads the same
    // texels; clamping it first keeps the float-to-int conversion defined.
    int ix = (int)fminf(fmaxf(bx, -4.0f), (float)(W + 4)), iy = (int)fminf(fmaxf(by, -4.0f), (float)(H + 4));
    for (int j = 0; j < 4; j++) {
        for (int k = 0; k < 4; k++) {
            int tx = min(max(ix + k - 1, 0), W - 1), ty = min(max(iy + j - 1, 0), H - 1);
            float4 h = hist[(size_t)ty * W + tx];
            float w = wx[k] * wy[j];
            r = f3{r.x + h.x * w, r.y + h.y * w, r.z + h.z * w};
        }
    }
    return f3{fmaxf(r.x, 0.0f), fmaxf(r.y, 0.0f), fmaxf(r.z, 0.0f)};
}

__device__ __forceinline__ f3 clipToBox(f3 history, f3 current, f3 bmin, f3 bmax)
{
#pragma clang fp contract(off)
    f3 d = f3{history.x - current.x, history.y - current.y, history.z - current.z};
    float dd[3] = {d.x, d.y, d.z}, cc[3] = {current.x, current.y, current.z};
    float lo[3] = {bmin.x, bmin.y, bmin.z}, hi[3] = {bmax.x, bmax.y, bmax.z};
    float tm = INFINITY;
    for (int k = 0; k < 3; k++) {
        float mag = fabsf(dd[k]);
        float lim = dd[k] > 0.0f ? hi[k] - cc[k] : cc[k] - lo[k];
        float t = mag > 1e-6f ? lim / fmaxf(mag, 1e-6f) : 1.0f;
        tm = fminf(tm, t);
    }
    float s = fminf(fmaxf(tm, 0.0f), 1.0f);
    return f3{current.x + d.x * s, current.y + d.y * s, current.z + d.z * s};
}

__global__ __launch_bounds__(128) void kTemporalResolve(const float4* accum, const float* depth, const float4* motion,
                                                       const float4* histIn, float4* histOut, TaaParams p)
{
#pragma clang fp contract(off)
    const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 8 + (threadIdx.x >> 4);
    const int W = (int)p.W, H = (int)p.H;
    if (x >= W || y >= H) return;
    const size_t c = (size_t)y * W + x;
    const float4 cur = accum[c];
    const f3 curY = toYCoCg(f3{cur.x, cur.y, cur.z});
    f3 m1 = splat(0.0f), m2 = splat(0.0f), nmin = splat(1e30f), nmax = splat(-1e30f);
    for (int j = -1; j <= 1; j++) {
        for (int i = -1; i <= 1; i++) {
            f3 v;
            if (i == 0 && j == 0) {
                v = curY;
            } else {
                int nx = min(max(x + i, 0), W - 1), ny = min(max(y + j, 0), H - 1);
                float4 a = accum[(size_t)ny * W + nx];
                v = toYCoCg(f3{a.x, a.y, a.z});
            }
            m1 = f3{m1.x + v.x, m1.y + v.y, m1.z + v.z};
            m2 = f3{m2.x + v.x * v.x, m2.y + v.y * v.y, m2.z + v.z * v.z};
            nmin = f3{fminf(nmin.x, v.x), fminf(nmin.y, v.y), fminf(nmin.z, v.z)};
            nmax = f3{fmaxf(nmax.x, v.x), fmaxf(nmax.y, v.y), fmaxf(nmax.z, v.z)};
        }
    }
    f3 mean = f3{m1.x / 9.0f, m1.y / 9.0f, m1.z / 9.0f};
    f3 sig = f3{sqrtf(fmaxf(m2.x / 9.0f - mean.x * mean.x, 0.0f)), sqrtf(fmaxf(m2.y / 9.0f - mean.y * mean.y, 0.0f)),
                sqrtf(fmaxf(m2.z / 9.0f - mean.z * mean.z, 0.0f))};
    f3 bmin = f3{fmaxf(mean.x - sig.x, nmin.x), fmaxf(mean.y - sig.y, nmin.y), fmaxf(mean.z - sig.z, nmin.z)};
    f3 bmax = f3{fminf(mean.x + sig.x, nmax.x), fminf(mean.y + sig.y, nmax.y), fminf(mean.z + sig.z, nmax.z)};
    bmin = f3{fminf(bmin.x, curY.x), fminf(bmin.y, curY.y), fminf(bmin.z, curY.z)};
    bmax = f3{fmaxf(bmax.x, curY.x), fmaxf(bmax.y, curY.y), fmaxf(bmax.z, curY.z)};
    const float4 mv = motion[c];
    float2 prev = make_float2(((float)x + 0.5f) + mv.x, ((float)y + 0.5f) + mv.y);
    bool valid = p.isReset == 0;
    valid = valid && prev.x >= 0.0f && prev.y >= 0.0f && prev.x < (float)W && prev.y < (float)H;
    float vz = fabsf(depth[c]);
    valid = valid && (fabsf(mv.z) < 0.05f * fmaxf(vz, 1e-4f));
    if (!(isfinite(prev.x) && isfinite(prev.y))) prev = make_float2(0.0f, 0.0f);  // already invalid
    f3 hist = catmullRom(histIn, prev, W, H);
    f3 hY = clipToBox(toYCoCg(hist), curY, bmin, bmax);
    hist = fromYCoCg(hY);
    valid = valid && isfinite(hist.x) && isfinite(hist.y) && isfinite(hist.z);
    float speed = sqrtf(mv.x * mv.x + mv.y * mv.y);
    float s = fminf(fmaxf(speed / fmaxf(p.speedReference, 1e-4f), 0.0f), 1.0f);
    float alpha = p.alphaMin + (p.alphaMax - p.alphaMin) * s;
    alpha = valid ? fminf(fmaxf(alpha, 0.0f), 1.0f) : 1.0f;
    histOut[c] = make_float4(hist.x + (cur.x - hist.x) * alpha, hist.y + (cur.y - hist.y) * alpha,
                             hist.z + (cur.z - hist.z) * alpha, cur.w);
}

}  // namespace

hipError_t launchAovs(const RenderScene& rs, const AovParams& p, const AovPlanes& out, hipStream_t st)
{
    uint32_t n = p.W * p.H;
    if (n == 0) return hipSuccess;
    kAovs<<<(n + kAovBlock - 1) / kAovBlock, kAovBlock, 0, st>>>(rs, p, out);
    return hipGetLastError();
}

hipError_t launchTemporalResolve(const float4* accum, const float* depth, const float4* motion, const float4* histIn,
                                 float4* histOut, const TaaParams& p, hipStream_t st)
{
    if (p.W == 0 || p.H == 0) return hipSuccess;
    dim3 grid((p.W + 15) / 16, (p.H + 7) / 8);
    kTemporalResolve<<<grid, 128, 0, st>>>(accum, depth, motion, histIn, histOut, p);
    return hipGetLastError();
}

}  // namespace aph
