// LBVH construction on gfx950 (Karras, "Maximizing Parallelism in the
// Construction of BVHs, Octrees, and k-d Trees", HPG 2012), used for both the
// per-mesh BLAS (triangles) and the TLAS (instance world boxes).
//
// Replaces the driver builds the reference requests with build flags NONE
// (HGIGeometry.cpp:127-160, PTGeometry.cpp:155-223 :185; HGIScene.cpp:377-420,
// PTScene.cpp:1460-1552 :1518).
//
// Passes (one launch each, all O(n)):
//   primBounds   per-primitive AABB + centroid, block-reduced centroid bounds
//   morton       63-bit Morton key of the normalized centroid
//   radix sort   (key, primitive) pairs
//   karras       internal node children + parents from the sorted keys
//   refit        bottom-up boxes; the second child to arrive builds the parent
//   assemble     64-B two-child nodes with both child boxes in the parent
//   reorder      triangles gathered into leaf order (BLAS only)
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "bvh_build.h"
#include "scene.h"

namespace aph {

namespace {

constexpr int kB = 256;

__device__ __forceinline__ uint32_t fOrd(float f)
{
    uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float fUnord(uint32_t u)
{
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

// Triangle AABBs from indexed positions.
__global__ void kTriBoxes(const float* pos, const uint32_t* idx, uint32_t n, float4* lo, float4* hi)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t a = idx[3 * i], b = idx[3 * i + 1], c = idx[3 * i + 2];
    float ax = pos[3 * a], ay = pos[3 * a + 1], az = pos[3 * a + 2];
    float bx = pos[3 * b], by = pos[3 * b + 1], bz = pos[3 * b + 2];
    float cx = pos[3 * c], cy = pos[3 * c + 1], cz = pos[3 * c + 2];
    lo[i] = make_float4(fminf(ax, fminf(bx, cx)), fminf(ay, fminf(by, cy)), fminf(az, fminf(bz, cz)), 0.0f);
    hi[i] = make_float4(fmaxf(ax, fmaxf(bx, cx)), fmaxf(ay, fmaxf(by, cy)), fmaxf(az, fmaxf(bz, cz)), 0.0f);
}

// Centroid bounds: ordered-uint atomics on 6 words (min xyz, max xyz).
__global__ void kCentroidBounds(const float4* lo, const float4* hi, uint32_t n, uint32_t* out6)
{
    __shared__ uint32_t s[6];
    if (threadIdx.x < 6) s[threadIdx.x] = threadIdx.x < 3 ? 0xFFFFFFFFu : 0u;
    __syncthreads();
    uint32_t mn[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, mx[3] = {0u, 0u, 0u};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float4 l = lo[i], h = hi[i];
        float c[3] = {0.5f * (l.x + h.x), 0.5f * (l.y + h.y), 0.5f * (l.z + h.z)};
#pragma unroll
        for (int k = 0; k < 3; k++) {
            uint32_t o = fOrd(c[k]);
            mn[k] = min(mn[k], o);
            mx[k] = max(mx[k], o);
        }
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
        atomicMin(&s[k], mn[k]);
        atomicMax(&s[3 + k], mx[k]);
    }
    __syncthreads();
    if (threadIdx.x < 3) atomicMin(&out6[threadIdx.x], s[threadIdx.x]);
    else if (threadIdx.x < 6) atomicMax(&out6[threadIdx.x], s[threadIdx.x]);
}

__device__ __forceinline__ uint64_t spread21(uint64_t v)
{
    v &= 0x1FFFFFull;
    v = (v | (v << 32)) & 0x1F00000000FFFFull;
    v = (v | (v << 16)) & 0x1F0000FF0000FFull;
    v = (v | (v << 8)) & 0x100F00F00F00F00Full;
    v = (v | (v << 4)) & 0x10C30C30C30C30C3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

__global__ void kMorton(const float4* lo, const float4* hi, uint32_t n, const uint32_t* bounds6, uint64_t* keys,
                        uint32_t* vals)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float bmin[3], inv[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        bmin[k] = fUnord(bounds6[k]);
        float ext = fUnord(bounds6[3 + k]) - bmin[k];
        inv[k] = ext > 0.0f ? 1.0f / ext : 0.0f;
    }
    float4 l = lo[i], h = hi[i];
    float c[3] = {0.5f * (l.x + h.x), 0.5f * (l.y + h.y), 0.5f * (l.z + h.z)};
    uint64_t q[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        float t = (c[k] - bmin[k]) * inv[k];
        t = fminf(fmaxf(t, 0.0f), 1.0f);
        q[k] = (uint64_t)fminf(t * 2097152.0f, 2097151.0f);
    }
    keys[i] = (spread21(q[0]) << 2) | (spread21(q[1]) << 1) | spread21(q[2]);
    vals[i] = i;
}

__device__ __forceinline__ int delta(const uint64_t* keys, int n, int i, int j)
{
    if (j < 0 || j >= n) return -1;
    uint64_t a = keys[i], b = keys[j];
    if (a == b) return 64 + __clz((uint32_t)(i ^ j));
    return __clzll((long long)(a ^ b));
}

// Internal node i in [0, n-2]; node references: internal k -> k, leaf k -> (n-1)+k in parent[].
__global__ void kKarras(const uint64_t* keys, int n, int2* child, uint32_t* parent)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = delta(keys, n, i, i - d);
    int lmax = 2;
    while (delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int dnode = delta(keys, n, i, j);
    int s = 0, t = l;
    do {
        t = (t + 1) >> 1;
        if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    int gamma = i + s * d + min(d, 0);
    int lo = min(i, j), hi = max(i, j);
    // encode: >= 0 internal, < 0 leaf (-(leaf+1))
    int left = (lo == gamma) ? -(gamma + 1) : gamma;
    int right = (hi == gamma + 1) ? -(gamma + 2) : gamma + 1;
    child[i] = make_int2(left, right);
    parent[left >= 0 ? left : (n - 1) + (-left - 1)] = (uint32_t)i;
    parent[right >= 0 ? right : (n - 1) + (-right - 1)] = (uint32_t)i;
}

// Bottom-up boxes.  boxLo/boxHi are indexed like parent[] (internal then leaf).
__global__ void kRefit(const float4* plo, const float4* phi, const uint32_t* sorted, int n, const int2* child,
                       const uint32_t* parent, uint32_t* visits, float4* boxLo, float4* boxHi)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t p = sorted[i];
    boxLo[(n - 1) + i] = plo[p];
    boxHi[(n - 1) + i] = phi[p];
    if (n == 1) return;
    uint32_t node = parent[(n - 1) + i];
    __threadfence();
    for (int guard = 0; guard < n; guard++) {
        if (atomicAdd(&visits[node], 1u) == 0u) return;  // first child to arrive stops
        __threadfence();
        int2 c = child[node];
        int a = c.x >= 0 ? c.x : (n - 1) + (-c.x - 1);
        int b = c.y >= 0 ? c.y : (n - 1) + (-c.y - 1);
        float4 la = boxLo[a], ha = boxHi[a];
        float4 lb = boxLo[b], hb = boxHi[b];
        boxLo[node] = make_float4(fminf(la.x, lb.x), fminf(la.y, lb.y), fminf(la.z, lb.z), 0.0f);
        boxHi[node] = make_float4(fmaxf(ha.x, hb.x), fmaxf(ha.y, hb.y), fmaxf(ha.z, hb.z), 0.0f);
        __threadfence();
        if (node == 0) return;
        node = parent[node];
    }
}

__global__ void kAssemble(int n, const int2* child, const float4* boxLo, const float4* boxHi, const uint32_t* sorted,
                          int leafOriginal, const uint32_t* remap, BvhNode* nodes)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    int2 c = child[i];
    int ch[2] = {c.x, c.y};
    int32_t ref[2];
    float4 lo[2], hi[2];
#pragma unroll
    for (int k = 0; k < 2; k++) {
        int slot;
        if (ch[k] >= 0) {
            ref[k] = ch[k];
            slot = ch[k];
        } else {
            uint32_t leaf = (uint32_t)(-ch[k] - 1);
            uint32_t prim = leafOriginal ? (remap ? remap[sorted[leaf]] : sorted[leaf]) : leaf;
            ref[k] = makeLeaf(prim, 1);
            slot = (n - 1) + (int)leaf;
        }
        lo[k] = boxLo[slot];
        hi[k] = boxHi[slot];
    }
    BvhNode nd;
    nd.a = make_float4(lo[0].x, lo[0].y, lo[0].z, __int_as_float(ref[0]));
    nd.b = make_float4(hi[0].x, hi[0].y, hi[0].z, __int_as_float(ref[1]));
    nd.c = make_float4(lo[1].x, lo[1].y, lo[1].z, 0.0f);
    nd.d = make_float4(hi[1].x, hi[1].y, hi[1].z, 0.0f);
    nodes[i] = nd;
}

__global__ void kSingleLeaf(const float4* plo, const float4* phi, uint32_t leafRef, BvhNode* nodes)
{
    float4 lo = plo[0], hi = phi[0];
    const float e = 3.0e38f;  // inverted (empty) box: no ray has tnear <= tfar
    BvhNode nd;
    nd.a = make_float4(lo.x, lo.y, lo.z, __int_as_float((int32_t)leafRef));
    nd.b = make_float4(hi.x, hi.y, hi.z, __int_as_float((int32_t)leafRef));
    nd.c = make_float4(e, e, e, 0.0f);
    nd.d = make_float4(-e, -e, -e, 0.0f);
    nodes[0] = nd;
}

__global__ void kDepth(int n, const uint32_t* parent, uint32_t* maxDepth)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || n == 1) return;
    uint32_t node = parent[(n - 1) + i], d = 1;
    while (node != 0 && d <= (uint32_t)n) {
        node = parent[node];
        d++;
    }
    atomicMax(maxDepth, d);
}

__global__ void kReorderTris(const float* pos, const uint32_t* idx, const uint32_t* sorted, uint32_t n, Tri* tris)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t p = sorted[i];
    uint32_t a = idx[3 * p], b = idx[3 * p + 1], c = idx[3 * p + 2];
    Tri t;
    t.v0 = make_float4(pos[3 * a], pos[3 * a + 1], pos[3 * a + 2], __uint_as_float(p));
    t.v1 = make_float4(pos[3 * b], pos[3 * b + 1], pos[3 * b + 2], 0.0f);
    t.v2 = make_float4(pos[3 * c], pos[3 * c + 1], pos[3 * c + 2], 0.0f);
    tris[i] = t;
}

// Four-child copy of node i: each internal child is replaced by its own two
// children (boxes and references copied bit for bit), leaves stay; unused slots
// get an inverted box, which every slab test rejects, and kEmptyChild.
__global__ void kCollapse4(const BvhNode* bin, uint32_t n, Bvh4Node* out)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float e = 3.0e38f;
    float lo[3][4], hi[3][4];
    int32_t ref[4];
    int k = 0;
    auto add = [&](float4 l, float4 h, int32_t r) {
        lo[0][k] = l.x; lo[1][k] = l.y; lo[2][k] = l.z;
        hi[0][k] = h.x; hi[1][k] = h.y; hi[2][k] = h.z;
        ref[k] = r;
        k++;
    };
    const BvhNode nd = bin[i];
    const float4 cl[2] = {nd.a, nd.c}, ch[2] = {nd.b, nd.d};
    const int32_t cr[2] = {__float_as_int(nd.a.w), __float_as_int(nd.b.w)};
    for (int j = 0; j < 2; j++) {
        if (cr[j] >= 0) {
            const BvhNode g = bin[cr[j]];
            add(g.a, g.b, __float_as_int(g.a.w));
            add(g.c, g.d, __float_as_int(g.b.w));
        } else {
            add(cl[j], ch[j], cr[j]);
        }
    }
    while (k < 4) add(make_float4(e, e, e, 0.0f), make_float4(-e, -e, -e, 0.0f), kEmptyChild);
    Bvh4Node o;
    o.lox = make_float4(lo[0][0], lo[0][1], lo[0][2], lo[0][3]);
    o.loy = make_float4(lo[1][0], lo[1][1], lo[1][2], lo[1][3]);
    o.loz = make_float4(lo[2][0], lo[2][1], lo[2][2], lo[2][3]);
    o.hix = make_float4(hi[0][0], hi[0][1], hi[0][2], hi[0][3]);
    o.hiy = make_float4(hi[1][0], hi[1][1], hi[1][2], hi[1][3]);
    o.hiz = make_float4(hi[2][0], hi[2][1], hi[2][2], hi[2][3]);
    o.ref = make_int4(ref[0], ref[1], ref[2], ref[3]);
    o.pad = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    out[i] = o;
}

inline unsigned blocks(uint32_t n) { return (n + kB - 1) / kB; }

#define APH_CHECK(x)                                  \
    do {                                              \
        hipError_t e_ = (x);                          \
        if (e_ != hipSuccess) return e_;              \
    } while (0)

}  // namespace

hipError_t lbvhBuild(const float4* primLo, const float4* primHi, uint32_t n, bool leafOriginal,
                     const uint32_t* leafRemap, hipStream_t st, BvhNode* nodes, uint32_t* sortedOut, BuildInfo* info)
{
    info->depth = 0;
    info->nodes = 0;
    if (n == 0) return hipSuccess;
    if (n == 1) {
        uint32_t prim = 0;
        if (leafOriginal && leafRemap) {
            hipError_t e = hipMemcpyAsync(&prim, leafRemap, sizeof(uint32_t), hipMemcpyDeviceToHost, st);
            if (e == hipSuccess) e = hipStreamSynchronize(st);
            if (e != hipSuccess) return e;
        }
        kSingleLeaf<<<1, 1, 0, st>>>(primLo, primHi, (uint32_t)makeLeaf(prim, 1), nodes);
        APH_CHECK(hipGetLastError());
        uint32_t zero = 0;
        APH_CHECK(hipMemcpyAsync(sortedOut, &zero, sizeof(uint32_t), hipMemcpyHostToDevice, st));
        float4 lo, hi;
        APH_CHECK(hipMemcpyAsync(&lo, primLo, sizeof(float4), hipMemcpyDeviceToHost, st));
        APH_CHECK(hipMemcpyAsync(&hi, primHi, sizeof(float4), hipMemcpyDeviceToHost, st));
        APH_CHECK(hipStreamSynchronize(st));
        info->rootLo[0] = lo.x; info->rootLo[1] = lo.y; info->rootLo[2] = lo.z;
        info->rootHi[0] = hi.x; info->rootHi[1] = hi.y; info->rootHi[2] = hi.z;
        info->depth = 1;
        info->nodes = 1;
        return hipSuccess;
    }
    // scratch
    uint32_t* bounds6 = nullptr;
    uint64_t *keys = nullptr, *keys2 = nullptr;
    uint32_t *vals = nullptr, *parent = nullptr, *visits = nullptr, *depth = nullptr;
    int2* child = nullptr;
    float4 *boxLo = nullptr, *boxHi = nullptr;
    void* sortTmp = nullptr;
    size_t sortBytes = 0;
    hipError_t err = hipSuccess;
    do {
        if ((err = hipMalloc(&bounds6, 8 * sizeof(uint32_t))) != hipSuccess) break;
        if ((err = hipMalloc(&keys, n * sizeof(uint64_t))) != hipSuccess) break;
        if ((err = hipMalloc(&keys2, n * sizeof(uint64_t))) != hipSuccess) break;
        if ((err = hipMalloc(&vals, n * sizeof(uint32_t))) != hipSuccess) break;
        if ((err = hipMalloc(&parent, (2 * (size_t)n) * sizeof(uint32_t))) != hipSuccess) break;
        if ((err = hipMalloc(&visits, n * sizeof(uint32_t))) != hipSuccess) break;
        if ((err = hipMalloc(&depth, sizeof(uint32_t))) != hipSuccess) break;
        if ((err = hipMalloc(&child, n * sizeof(int2))) != hipSuccess) break;
        if ((err = hipMalloc(&boxLo, 2 * (size_t)n * sizeof(float4))) != hipSuccess) break;
        if ((err = hipMalloc(&boxHi, 2 * (size_t)n * sizeof(float4))) != hipSuccess) break;
        uint32_t init6[8] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u, 0u, 0u, 0u};
        if ((err = hipMemcpyAsync(bounds6, init6, sizeof(init6), hipMemcpyHostToDevice, st)) != hipSuccess) break;
        if ((err = hipMemsetAsync(visits, 0, n * sizeof(uint32_t), st)) != hipSuccess) break;
        if ((err = hipMemsetAsync(depth, 0, sizeof(uint32_t), st)) != hipSuccess) break;
        unsigned rb = min(blocks(n), 1024u);
        kCentroidBounds<<<rb, kB, 0, st>>>(primLo, primHi, n, bounds6);
        kMorton<<<blocks(n), kB, 0, st>>>(primLo, primHi, n, bounds6, keys, vals);
        if ((err = hipGetLastError()) != hipSuccess) break;
        if ((err = hipcub::DeviceRadixSort::SortPairs(nullptr, sortBytes, keys, keys2, vals, sortedOut, (int)n, 0, 63,
                                                       st)) != hipSuccess)
            break;
        if ((err = hipMalloc(&sortTmp, sortBytes)) != hipSuccess) break;
        if ((err = hipcub::DeviceRadixSort::SortPairs(sortTmp, sortBytes, keys, keys2, vals, sortedOut, (int)n, 0, 63,
                                                       st)) != hipSuccess)
            break;
        kKarras<<<blocks(n - 1), kB, 0, st>>>(keys2, (int)n, child, parent);
        kRefit<<<blocks(n), kB, 0, st>>>(primLo, primHi, sortedOut, (int)n, child, parent, visits, boxLo, boxHi);
        kAssemble<<<blocks(n - 1), kB, 0, st>>>((int)n, child, boxLo, boxHi, sortedOut, leafOriginal ? 1 : 0, leafRemap,
                                                nodes);
        kDepth<<<blocks(n), kB, 0, st>>>((int)n, parent, depth);
        if ((err = hipGetLastError()) != hipSuccess) break;
        float4 rlo, rhi;
        uint32_t dep = 0;
        if ((err = hipMemcpyAsync(&rlo, boxLo, sizeof(float4), hipMemcpyDeviceToHost, st)) != hipSuccess) break;
        if ((err = hipMemcpyAsync(&rhi, boxHi, sizeof(float4), hipMemcpyDeviceToHost, st)) != hipSuccess) break;
        if ((err = hipMemcpyAsync(&dep, depth, sizeof(uint32_t), hipMemcpyDeviceToHost, st)) != hipSuccess) break;
        if ((err = hipStreamSynchronize(st)) != hipSuccess) break;
        info->rootLo[0] = rlo.x; info->rootLo[1] = rlo.y; info->rootLo[2] = rlo.z;
        info->rootHi[0] = rhi.x; info->rootHi[1] = rhi.y; info->rootHi[2] = rhi.z;
        info->depth = dep;
        info->nodes = n - 1;
    } while (0);
    hipFree(bounds6); hipFree(keys); hipFree(keys2); hipFree(vals); hipFree(parent); hipFree(visits);
    hipFree(depth); hipFree(child); hipFree(boxLo); hipFree(boxHi); hipFree(sortTmp);
    return err;
}

hipError_t collapse4(const BvhNode* nodes, uint32_t n, hipStream_t st, Bvh4Node* out)
{
    if (n == 0) return hipSuccess;
    kCollapse4<<<blocks(n), kB, 0, st>>>(nodes, n, out);
    return hipGetLastError();
}

hipError_t triBoxes(const float* pos, const uint32_t* idx, uint32_t n, hipStream_t st, float4* lo, float4* hi)
{
    if (n == 0) return hipSuccess;
    kTriBoxes<<<blocks(n), kB, 0, st>>>(pos, idx, n, lo, hi);
    return hipGetLastError();
}

hipError_t reorderTris(const float* pos, const uint32_t* idx, const uint32_t* sorted, uint32_t n, hipStream_t st,
                       Tri* tris)
{
    if (n == 0) return hipSuccess;
    kReorderTris<<<blocks(n), kB, 0, st>>>(pos, idx, sorted, n, tris);
    return hipGetLastError();
}

}  // namespace aph
