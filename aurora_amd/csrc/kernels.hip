// Ray-query, path-tracing and post-processing kernels for gfx950.
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "pathtrace.h"

namespace aph {

namespace {

constexpr int kBlock = 128;  // 2 waves of 64; 16x8 pixel tiles for the megakernel
static_assert(kBlock == kStackLanes, "the traversal stack stride is the block size");

__device__ __forceinline__ uint32_t waveSum(uint32_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Flushes per-lane counters with one atomic per wave; every lane of the wave must call it.
__device__ __forceinline__ void flushCounters(const PathCounters& pc, uint32_t paths, bool stats, DevCounters* out)
{
    uint32_t ext = waveSum(pc.ext), sh = waveSum(pc.shadow), np = waveSum(paths);
    uint32_t nodes = 0, tris = 0, insts = 0, snodes = 0, stris = 0, sinsts = 0;
    if (stats) {
        snodes = waveSum(pc.tsS.nodes);
        stris = waveSum(pc.tsS.tris);
        sinsts = waveSum(pc.tsS.insts);
        nodes = waveSum(pc.ts.nodes) + snodes;
        tris = waveSum(pc.ts.tris) + stris;
        insts = waveSum(pc.ts.insts) + sinsts;
    }
    uint32_t ovf = waveSum(pc.ts.overflow | pc.tsS.overflow);
    if ((threadIdx.x & 63) == 0) {
        if (ext) atomicAdd(&out->ext, (unsigned long long)ext);
        if (sh) atomicAdd(&out->shadow, (unsigned long long)sh);
        if (np) atomicAdd(&out->paths, (unsigned long long)np);
        if (nodes) atomicAdd(&out->nodes, (unsigned long long)nodes);
        if (tris) atomicAdd(&out->tris, (unsigned long long)tris);
        if (insts) atomicAdd(&out->insts, (unsigned long long)insts);
        if (snodes) atomicAdd(&out->snodes, (unsigned long long)snodes);
        if (stris) atomicAdd(&out->stris, (unsigned long long)stris);
        if (sinsts) atomicAdd(&out->sinsts, (unsigned long long)sinsts);
        if (ovf) atomicAdd(&out->overflow, (unsigned long long)ovf);
    }
}

// Megakernel: one lane per pixel, samples [start, start+count) in order, running
// mean accumulated exactly like Accumulation.glsl:22-30.
template <bool STATS, bool GROUND>
__global__ __launch_bounds__(kBlock) void kRenderMega(RenderScene rs, aph_frame fp, uint32_t W, uint32_t H,
                                                     uint32_t start, uint32_t count, uint32_t seedOffset,
                                                     float4* accum, DevCounters* ctr)
{
    __shared__ uint32_t sStack[kLdsStack * kBlock];
    uint32_t spill[kSpillStack];
    const StackMem stk{sStack + threadIdx.x, spill};
    PathCounters pc = {};
    // 16x8 tile per block, row-major tiles
    uint32_t tilesX = (W + 15) / 16;
    uint32_t tx = blockIdx.x % tilesX, ty = blockIdx.x / tilesX;
    uint32_t px = tx * 16 + (threadIdx.x & 15), py = ty * 8 + (threadIdx.x >> 4);
    uint32_t paths = 0;
    if (px < W && py < H) {
        size_t pix = (size_t)py * W + px;
        float4 acc = accum[pix];
        for (uint32_t s = start; s < start + count; s++) {
            float4 r = pathSample<STATS, GROUND>(rs, fp, W, H, px, py, s + seedOffset, stk, pc);
            if (s > 0) {
                float t = 1.0f / (float)(s + 1);
                float u = 1.0f - t;
                acc = make_float4(acc.x * u + r.x * t, acc.y * u + r.y * t, acc.z * u + r.z * t, acc.w * u + r.w * t);
            } else {
                acc = r;
            }
            paths++;
        }
        accum[pix] = acc;
    }
    flushCounters(pc, paths, STATS, ctr);
}

template <bool ANY, bool STATS>
__global__ __launch_bounds__(kBlock) void kTrace(TraceScene sc, const aph_ray* rays, uint32_t n, aph_hit* hits,
                                                uint8_t* occ, DevCounters* ctr)
{
    __shared__ uint32_t sStack[kLdsStack * kBlock];
    uint32_t spill[kSpillStack];
    const StackMem stk{sStack + threadIdx.x, spill};
    PathCounters pc = {};
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        aph_ray r = rays[i];
        HitRec h;
        bool found = traceRay<ANY, STATS>(sc, f3{r.org[0], r.org[1], r.org[2]}, f3{r.dir[0], r.dir[1], r.dir[2]},
                                          r.tmin, r.tmax, stk, h, ANY ? pc.tsS : pc.ts);
        if (ANY) {
            occ[i] = found ? 1 : 0;
            pc.shadow = 1;
        } else {
            aph_hit o;
            if (found) {
                o.t = (float)h.t; o.inst = h.inst; o.prim = h.prim; o.u = (float)h.u; o.v = (float)h.v;
            } else {
                o.t = INFINITY; o.inst = 0xFFFFFFFFu; o.prim = 0xFFFFFFFFu; o.u = 0.0f; o.v = 0.0f;
            }
            hits[i] = o;
            pc.ext = 1;
        }
    }
    flushCounters(pc, 0, STATS, ctr);
}

// PostProcessing.glsl:37-64 and Colors.slang:33-53.
__global__ void kPost(const float4* accum, uint32_t n, float3 brightness, int toneMap, int gamma, float4* out)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 a = accum[i];
    f3 c = f3{a.x * brightness.x, a.y * brightness.y, a.z * brightness.z};
    if (toneMap) {
        f3 num = c * (c * 2.51f + splat(0.03f));
        f3 den = c * (c * 2.43f + splat(0.59f)) + splat(0.14f);
        f3 r = num / den;
        c = f3{sat(r.x), sat(r.y), sat(r.z)};
    }
    if (gamma) {
        c = f3{sat(c.x), sat(c.y), sat(c.z)};
        f3 s1 = f3{sqrtf(c.x), sqrtf(c.y), sqrtf(c.z)};
        f3 s2 = f3{sqrtf(s1.x), sqrtf(s1.y), sqrtf(s1.z)};
        f3 s3 = f3{sqrtf(s2.x), sqrtf(s2.y), sqrtf(s2.z)};
        c = s1 * 0.662002687f + s2 * 0.684122060f - (s3 * 0.323583601f + c * 0.0225411470f);
    }
    out[i] = make_float4(c.x, c.y, c.z, 1.0f);
}

}  // namespace

hipError_t launchRenderMega(const RenderScene& rs, const aph_frame& fp, uint32_t W, uint32_t H, uint32_t start,
                            uint32_t count, uint32_t seedOffset, float4* accum, DevCounters* ctr, bool stats,
                            hipStream_t st)
{
    uint32_t tiles = ((W + 15) / 16) * ((H + 7) / 8);
    if (tiles == 0 || count == 0) return hipSuccess;
    const bool ground = groundActive(rs.ground);
    if (stats && ground)
        kRenderMega<true, true><<<tiles, kBlock, 0, st>>>(rs, fp, W, H, start, count, seedOffset, accum, ctr);
    else if (stats)
        kRenderMega<true, false><<<tiles, kBlock, 0, st>>>(rs, fp, W, H, start, count, seedOffset, accum, ctr);
    else if (ground)
        kRenderMega<false, true><<<tiles, kBlock, 0, st>>>(rs, fp, W, H, start, count, seedOffset, accum, ctr);
    else
        kRenderMega<false, false><<<tiles, kBlock, 0, st>>>(rs, fp, W, H, start, count, seedOffset, accum, ctr);
    return hipGetLastError();
}

hipError_t launchTrace(const TraceScene& sc, const aph_ray* rays, uint32_t n, aph_hit* hits, uint8_t* occ,
                       DevCounters* ctr, bool any, bool stats, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    uint32_t blocks = (n + kBlock - 1) / kBlock;
    if (any) {
        if (stats) kTrace<true, true><<<blocks, kBlock, 0, st>>>(sc, rays, n, hits, occ, ctr);
        else kTrace<true, false><<<blocks, kBlock, 0, st>>>(sc, rays, n, hits, occ, ctr);
    } else {
        if (stats) kTrace<false, true><<<blocks, kBlock, 0, st>>>(sc, rays, n, hits, occ, ctr);
        else kTrace<false, false><<<blocks, kBlock, 0, st>>>(sc, rays, n, hits, occ, ctr);
    }
    return hipGetLastError();
}

hipError_t launchPost(const float4* accum, uint32_t n, const float brightness[3], int toneMap, int gamma, float4* out,
                      hipStream_t st)
{
    if (n == 0) return hipSuccess;
    kPost<<<(n + 255) / 256, 256, 0, st>>>(accum, n, make_float3(brightness[0], brightness[1], brightness[2]),
                                           toneMap, gamma, out);
    return hipGetLastError();
}

}  // namespace aph
