// Binned-SAH BVH builder for gfx950 (APH_BUILDER_SAH).
//
// Replaces the driver's BLAS build behind PTGeometry::buildBLAS /
// HGIGeometry::update (PTGeometry.cpp:155-223, HGIGeometry.cpp:127-160), whose
// internals the reference does not contain; the output is the same two-child
// node format the LBVH builder writes, so traversal is unchanged.
//
// Top-down and level-synchronous.  Every task is a contiguous range of the
// primitive permutation with its box and centroid bounds; processing a task
// either makes it a leaf or splits it at the best of 3 x (kBins - 1) binned
// SAH planes (Wald, "On fast construction of SAH-based bounding volume
// hierarchies", RT 2007) into two child tasks of the next level.
//   * large tasks (> kSmallMax primitives): chunks of kChunk primitives bin
//     into LDS and flush to per-task global bins; one wave per task picks the
//     split; a scan over per-chunk counts gives every chunk its output offset,
//     so the partition is deterministic;
//   * small tasks: one wave per task bins in LDS, picks the split and
//     partitions with wave ballots, in one kernel.
// Child bounds come exactly from the bins (box and centroid bounds per bin),
// so no level re-reduces its primitives.  When all centroids of a task
// coincide, binning cannot separate them: the task splits its range in halves
// and both children take the parent's bounds (conservative, still containing).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <vector>

#include "bvh_build.h"
#include "scene.h"

namespace aph {

namespace {

constexpr int kBins = 16;
constexpr int kCand = 3 * (kBins - 1);  // split candidates per task
constexpr uint32_t kSmallMax = 256;     // tasks up to this size take the one-wave path
constexpr uint32_t kChunk = 2048;       // primitives per block in the large-task kernels
constexpr int kLB = 256;                // threads per large-task block
// Per bin and axis: count, box lo/hi, centroid lo/hi (ordered uints), 13 words.
constexpr int kBinWords = 13;
constexpr int kTaskBinWords = 3 * kBins * kBinWords;
constexpr float kCostNode = 1.0f;       // SAH: relative cost of a node visit ...
constexpr float kCostTri = 1.2f;        // ... and of a triangle test
// From this level on, tasks split their range in halves: the tree depth stays
// below kForceLevel + log2(n) + 1, inside the traversal stack.
constexpr uint32_t kForceLevel = 32;

struct Task {
    uint32_t begin, count, parent, slot;  // parent = 0xFFFFFFFF for the root
    float bLo[3], bHi[3];                 // box bounds of the primitives
    float cLo[3], cHi[3];                 // centroid bounds
};

struct Split {
    int axis, bin;     // left = bins [0, bin] on axis; axis -1: halves of the range; -2: leaf
    uint32_t nLeft;
    uint32_t node;     // node index of a split task
};

__device__ __forceinline__ uint32_t fOrd(float f)
{
    uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float fUnord(uint32_t u)
{
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

__device__ __forceinline__ float area(const float lo[3], const float hi[3])
{
    float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    if (!(dx >= 0.0f && dy >= 0.0f && dz >= 0.0f)) return 0.0f;
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}

// Bin of a centroid coordinate; the binning and partition kernels share it so
// that their counts agree exactly.
__device__ __forceinline__ int binOf(float c, float lo, float hi)
{
    float ext = hi - lo;
    if (!(ext > 0.0f)) return 0;
    int b = (int)((c - lo) * ((float)kBins / ext));
    return b < 0 ? 0 : (b >= kBins ? kBins - 1 : b);
}

__device__ __forceinline__ void binsInit(uint32_t* bins, int tid, int nthreads)
{
    for (int i = tid; i < kTaskBinWords; i += nthreads) {
        int w = i % kBinWords;
        // word 0 count; 1..3 box lo; 4..6 box hi; 7..9 centroid lo; 10..12 centroid hi
        bins[i] = (w == 0) ? 0u : ((w >= 1 && w <= 3) || (w >= 7 && w <= 9)) ? 0xFFFFFFFFu : 0u;
    }
}

__device__ __forceinline__ void binAdd(uint32_t* bins, const Task& t, float4 lo, float4 hi)
{
    float c[3] = {0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z)};
    float l[3] = {lo.x, lo.y, lo.z}, h[3] = {hi.x, hi.y, hi.z};
#pragma unroll
    for (int a = 0; a < 3; a++) {
        uint32_t* b = bins + (a * kBins + binOf(c[a], t.cLo[a], t.cHi[a])) * kBinWords;
        atomicAdd(&b[0], 1u);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            atomicMin(&b[1 + k], fOrd(l[k]));
            atomicMax(&b[4 + k], fOrd(h[k]));
            atomicMin(&b[7 + k], fOrd(c[k]));
            atomicMax(&b[10 + k], fOrd(c[k]));
        }
    }
}

// Union of bins [b0, b1] on axis a.
struct BinSum {
    uint32_t n;
    float bLo[3], bHi[3], cLo[3], cHi[3];
};
__device__ __forceinline__ BinSum binRange(const uint32_t* bins, int a, int b0, int b1)
{
    BinSum s;
    s.n = 0;
    uint32_t bl[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, bh[3] = {0u, 0u, 0u};
    uint32_t cl[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, ch[3] = {0u, 0u, 0u};
    for (int i = b0; i <= b1; i++) {
        const uint32_t* b = bins + (a * kBins + i) * kBinWords;
        if (b[0] == 0) continue;
        s.n += b[0];
        for (int k = 0; k < 3; k++) {
            bl[k] = min(bl[k], b[1 + k]);
            bh[k] = max(bh[k], b[4 + k]);
            cl[k] = min(cl[k], b[7 + k]);
            ch[k] = max(ch[k], b[10 + k]);
        }
    }
    for (int k = 0; k < 3; k++) {
        s.bLo[k] = fUnord(bl[k]);
        s.bHi[k] = fUnord(bh[k]);
        s.cLo[k] = fUnord(cl[k]);
        s.cHi[k] = fUnord(ch[k]);
    }
    return s;
}

// Best split of a task from its bins, evaluated by one wave (lane = candidate).
// A task of up to maxLeaf primitives may become a leaf (never the root).
__device__ Split chooseSplit(const uint32_t* bins, const Task& t, uint32_t maxLeaf, bool halves)
{
    const int lane = threadIdx.x & 63;
    float cost = INFINITY;
    if (lane < kCand) {
        int a = lane / (kBins - 1), k = lane % (kBins - 1);
        BinSum L = binRange(bins, a, 0, k), R = binRange(bins, a, k + 1, kBins - 1);
        if (L.n > 0 && R.n > 0)
            cost = area(L.bLo, L.bHi) * (float)L.n + area(R.bLo, R.bHi) * (float)R.n;
    }
    // wave argmin (ties to the lower candidate)
    float best = cost;
    int bestLane = lane;
    for (int off = 32; off > 0; off >>= 1) {
        float oc = __shfl_xor(best, off, 64);
        int ol = __shfl_xor(bestLane, off, 64);
        if (oc < best || (oc == best && ol < bestLane)) {
            best = oc;
            bestLane = ol;
        }
    }
    Split s;
    s.node = 0;
    float pa = area(t.bLo, t.bHi);
    float splitCost = pa > 0.0f ? kCostNode + kCostTri * best / pa : kCostNode + kCostTri * (float)t.count;
    float leafCost = kCostTri * (float)t.count;
    if (t.count <= 1 || (t.count <= maxLeaf && leafCost <= splitCost && t.parent != 0xFFFFFFFFu)) {
        s.axis = -2;
        s.bin = 0;
        s.nLeft = t.count;
    } else if (halves || !(best < INFINITY)) {
        // no plane separates the centroids
        if (t.count <= maxLeaf && t.parent != 0xFFFFFFFFu) {
            s.axis = -2;
            s.bin = 0;
            s.nLeft = t.count;
        } else {
            s.axis = -1;
            s.bin = 0;
            s.nLeft = t.count / 2;
        }
    } else {
        s.axis = bestLane / (kBins - 1);
        s.bin = bestLane % (kBins - 1);
        s.nLeft = binRange(bins, s.axis, 0, s.bin).n;
    }
    return s;
}

// Writes a task's box and reference into its parent's child slot (scalar
// stores: the two slots of a node are written by different tasks).
__device__ __forceinline__ void writeSlot(BvhNode* nodes, const Task& t, int32_t ref)
{
    float* nd = reinterpret_cast<float*>(nodes + t.parent);
    if (t.slot == 0) {
        nd[0] = t.bLo[0]; nd[1] = t.bLo[1]; nd[2] = t.bLo[2]; nd[3] = __int_as_float(ref);
        nd[4] = t.bHi[0]; nd[5] = t.bHi[1]; nd[6] = t.bHi[2];
    } else {
        nd[7] = __int_as_float(ref);
        nd[8] = t.bLo[0]; nd[9] = t.bLo[1]; nd[10] = t.bLo[2]; nd[11] = 0.0f;
        nd[12] = t.bHi[0]; nd[13] = t.bHi[1]; nd[14] = t.bHi[2]; nd[15] = 0.0f;
    }
}

struct Lists {
    Task* small;
    Task* large;
    uint32_t* counts;  // [0] small, [1] large, [2] node allocator, [3] max level
};

__device__ __forceinline__ void emitChild(Lists out, const Task& c)
{
    if (c.count > kSmallMax) out.large[atomicAdd(&out.counts[1], 1u)] = c;
    else out.small[atomicAdd(&out.counts[0], 1u)] = c;
}

// Finishes a decided task (one lane): leaf or node, parent slot, child tasks.
__device__ void finishTask(const Task& t, const Split& s, const uint32_t* bins, BvhNode* nodes, Lists out,
                           uint32_t nodeIdx, int leafOriginal, const uint32_t* remap, const uint32_t* idxIn,
                           uint32_t* sortedOut)
{
    if (s.axis == -2) {
        int32_t ref;
        if (leafOriginal) {
            uint32_t p = idxIn[t.begin];
            ref = makeLeaf(remap ? remap[p] : p, 1);
        } else {
            ref = makeLeaf(t.begin, t.count);
        }
        if (t.parent != 0xFFFFFFFFu) writeSlot(nodes, t, ref);
        return;
    }
    if (t.parent != 0xFFFFFFFFu) writeSlot(nodes, t, (int32_t)nodeIdx);
    Task l, r;
    l.begin = t.begin;
    l.count = s.nLeft;
    r.begin = t.begin + s.nLeft;
    r.count = t.count - s.nLeft;
    l.parent = r.parent = nodeIdx;
    l.slot = 0;
    r.slot = 1;
    if (s.axis >= 0) {
        BinSum L = binRange(bins, s.axis, 0, s.bin), R = binRange(bins, s.axis, s.bin + 1, kBins - 1);
        for (int k = 0; k < 3; k++) {
            l.bLo[k] = L.bLo[k]; l.bHi[k] = L.bHi[k]; l.cLo[k] = L.cLo[k]; l.cHi[k] = L.cHi[k];
            r.bLo[k] = R.bLo[k]; r.bHi[k] = R.bHi[k]; r.cLo[k] = R.cLo[k]; r.cHi[k] = R.cHi[k];
        }
    } else {
        for (int k = 0; k < 3; k++) {
            l.bLo[k] = r.bLo[k] = t.bLo[k];
            l.bHi[k] = r.bHi[k] = t.bHi[k];
            l.cLo[k] = r.cLo[k] = t.cLo[k];
            l.cHi[k] = r.cHi[k] = t.cHi[k];
        }
    }
    emitChild(out, l);
    emitChild(out, r);
}

__device__ __forceinline__ bool goesLeft(const Split& s, const Task& t, float4 lo, float4 hi, uint32_t pos)
{
    if (s.axis == -1) return pos < t.begin + s.nLeft;
    float c = 0.5f * (s.axis == 0 ? lo.x + hi.x : (s.axis == 1 ? lo.y + hi.y : lo.z + hi.z));
    return binOf(c, t.cLo[s.axis], t.cHi[s.axis]) <= s.bin;
}

// ---- root bounds ----
__global__ void kRootBounds(const float4* lo, const float4* hi, uint32_t n, uint32_t* out12)
{
    uint32_t m[12];
    for (int k = 0; k < 12; k++) m[k] = (k % 6) < 3 ? 0xFFFFFFFFu : 0u;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float4 l = lo[i], h = hi[i];
        float v[6] = {l.x, l.y, l.z, h.x, h.y, h.z};
        float c[3] = {0.5f * (l.x + h.x), 0.5f * (l.y + h.y), 0.5f * (l.z + h.z)};
        for (int k = 0; k < 3; k++) {
            m[k] = min(m[k], fOrd(v[k]));
            m[3 + k] = max(m[3 + k], fOrd(v[3 + k]));
            m[6 + k] = min(m[6 + k], fOrd(c[k]));
            m[9 + k] = max(m[9 + k], fOrd(c[k]));
        }
    }
    for (int k = 0; k < 12; k++) {
        if ((k % 6) < 3) atomicMin(&out12[k], m[k]);
        else atomicMax(&out12[k], m[k]);
    }
}

__global__ void kRootTask(const uint32_t* b12, uint32_t n, Task* small, Task* large, uint32_t* counts)
{
    Task t;
    t.begin = 0;
    t.count = n;
    t.parent = 0xFFFFFFFFu;
    t.slot = 0;
    for (int k = 0; k < 3; k++) {
        t.bLo[k] = fUnord(b12[k]);
        t.bHi[k] = fUnord(b12[3 + k]);
        t.cLo[k] = fUnord(b12[6 + k]);
        t.cHi[k] = fUnord(b12[9 + k]);
    }
    if (n > kSmallMax) { large[0] = t; counts[1] = 1; counts[0] = 0; }
    else { small[0] = t; counts[0] = 1; counts[1] = 0; }
    counts[2] = 0;
}

__global__ void kIota(uint32_t* v, uint32_t n)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = i;
}

// ---- small tasks: one wave per task ----
__global__ __launch_bounds__(64) void kSmall(const Task* tasks, const float4* plo, const float4* phi,
                                             const uint32_t* idxIn, uint32_t* idxOut, uint32_t* sortedOut,
                                             BvhNode* nodes, Lists out, uint32_t maxLeaf, int leafOriginal,
                                             const uint32_t* remap, int halves)
{
    __shared__ uint32_t bins[kTaskBinWords];
    __shared__ uint32_t sNode;
    const Task t = tasks[blockIdx.x];
    const int lane = threadIdx.x;
    binsInit(bins, lane, 64);
    __syncthreads();
    for (uint32_t i = lane; i < t.count; i += 64) {
        uint32_t p = idxIn[t.begin + i];
        binAdd(bins, t, plo[p], phi[p]);
    }
    __syncthreads();
    Split s = chooseSplit(bins, t, maxLeaf, halves != 0);
    if (s.axis == -2) {
        for (uint32_t i = lane; i < t.count; i += 64) sortedOut[t.begin + i] = idxIn[t.begin + i];
        if (lane == 0) finishTask(t, s, bins, nodes, out, 0, leafOriginal, remap, idxIn, sortedOut);
        return;
    }
    // partition in primitive order: left part first, right part after nLeft
    uint32_t nl = 0, nr = 0;
    for (uint32_t base = 0; base < t.count; base += 64) {
        uint32_t i = base + lane;
        bool valid = i < t.count;
        uint32_t p = valid ? idxIn[t.begin + i] : 0;
        bool left = valid && goesLeft(s, t, plo[p], phi[p], t.begin + i);
        uint64_t lm = __ballot(left), rm = __ballot(valid && !left);
        uint64_t below = (1ull << lane) - 1ull;
        if (left) idxOut[t.begin + nl + __popcll(lm & below)] = p;
        else if (valid) idxOut[t.begin + s.nLeft + nr + __popcll(rm & below)] = p;
        nl += __popcll(lm);
        nr += __popcll(rm);
    }
    if (lane == 0) {
        sNode = atomicAdd(&out.counts[2], 1u);
        finishTask(t, s, bins, nodes, out, sNode, leafOriginal, remap, idxIn, sortedOut);
    }
}

// ---- large tasks ----
struct Chunks {
    const uint32_t* task;   // task of each chunk
    const uint32_t* first;  // first chunk of each task
};

__global__ void kChunkCount(const Task* tasks, uint32_t m, uint32_t* cnt)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) cnt[i] = (tasks[i].count + kChunk - 1) / kChunk;
}

__global__ void kChunkList(const Task* tasks, uint32_t m, const uint32_t* first, uint32_t* chunkTask)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    uint32_t c = (tasks[i].count + kChunk - 1) / kChunk;
    for (uint32_t k = 0; k < c; k++) chunkTask[first[i] + k] = i;
}

__global__ void kBinsInit(uint32_t* bins, uint32_t m)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m * kTaskBinWords) return;
    int w = (int)(i % kBinWords);
    bins[i] = (w == 0) ? 0u : ((w >= 1 && w <= 3) || (w >= 7 && w <= 9)) ? 0xFFFFFFFFu : 0u;
}

__global__ __launch_bounds__(kLB) void kBinLarge(const Task* tasks, Chunks ch, const float4* plo, const float4* phi,
                                                 const uint32_t* idxIn, uint32_t* gbins)
{
    __shared__ uint32_t bins[kTaskBinWords];
    const uint32_t ti = ch.task[blockIdx.x];
    const Task t = tasks[ti];
    const uint32_t c0 = (blockIdx.x - ch.first[ti]) * kChunk;
    const uint32_t c1 = min(t.count, c0 + kChunk);
    binsInit(bins, threadIdx.x, kLB);
    __syncthreads();
    for (uint32_t i = c0 + threadIdx.x; i < c1; i += kLB) {
        uint32_t p = idxIn[t.begin + i];
        binAdd(bins, t, plo[p], phi[p]);
    }
    __syncthreads();
    uint32_t* g = gbins + (size_t)ti * kTaskBinWords;
    for (int i = threadIdx.x; i < kTaskBinWords; i += kLB) {
        int w = i % kBinWords;
        uint32_t v = bins[i];
        if (w == 0) { if (v) atomicAdd(&g[i], v); }
        else if ((w >= 1 && w <= 3) || (w >= 7 && w <= 9)) { if (v != 0xFFFFFFFFu) atomicMin(&g[i], v); }
        else if (v) atomicMax(&g[i], v);
    }
}

__global__ __launch_bounds__(64) void kSplitLarge(const Task* tasks, const uint32_t* gbins, Split* splits,
                                                  BvhNode* nodes, Lists out, int halves)
{
    const Task t = tasks[blockIdx.x];
    const uint32_t* bins = gbins + (size_t)blockIdx.x * kTaskBinWords;
    Split s = chooseSplit(bins, t, 0, halves != 0);
    if (threadIdx.x == 0) {
        s.node = atomicAdd(&out.counts[2], 1u);
        splits[blockIdx.x] = s;
        finishTask(t, s, bins, nodes, out, s.node, 0, nullptr, nullptr, nullptr);
    }
}

// Block-wide exclusive prefix of a flag over kLB threads.
__device__ __forceinline__ uint32_t blockPrefix(bool flag, uint32_t* sw, uint32_t& total)
{
    uint64_t b = __ballot(flag);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) sw[wave] = (uint32_t)__popcll(b);
    __syncthreads();
    uint32_t off = 0, tot = 0;
    for (int w = 0; w < kLB / 64; w++) {
        if (w < wave) off += sw[w];
        tot += sw[w];
    }
    __syncthreads();
    total = tot;
    return off + (uint32_t)__popcll(b & ((1ull << lane) - 1ull));
}

__global__ __launch_bounds__(kLB) void kPartCount(const Task* tasks, const Split* splits, Chunks ch,
                                                  const float4* plo, const float4* phi, const uint32_t* idxIn,
                                                  uint32_t* leftCnt)
{
    __shared__ uint32_t sw[kLB / 64];
    const uint32_t ti = ch.task[blockIdx.x];
    const Task t = tasks[ti];
    const Split s = splits[ti];
    const uint32_t c0 = (blockIdx.x - ch.first[ti]) * kChunk;
    const uint32_t c1 = min(t.count, c0 + kChunk);
    uint32_t nl = 0;
    for (uint32_t base = c0; base < c1; base += kLB) {
        uint32_t i = base + threadIdx.x;
        bool left = false;
        if (i < c1) {
            uint32_t p = idxIn[t.begin + i];
            left = goesLeft(s, t, plo[p], phi[p], t.begin + i);
        }
        uint32_t tot;
        blockPrefix(left, sw, tot);
        nl += tot;
    }
    if (threadIdx.x == 0) leftCnt[blockIdx.x] = nl;
}

__global__ __launch_bounds__(kLB) void kPartScatter(const Task* tasks, const Split* splits, Chunks ch,
                                                    const float4* plo, const float4* phi, const uint32_t* idxIn,
                                                    const uint32_t* leftScan, uint32_t* idxOut)
{
    __shared__ uint32_t sw[kLB / 64];
    const uint32_t ti = ch.task[blockIdx.x];
    const Task t = tasks[ti];
    const Split s = splits[ti];
    const uint32_t chunk = blockIdx.x - ch.first[ti];
    const uint32_t c0 = chunk * kChunk;
    const uint32_t c1 = min(t.count, c0 + kChunk);
    // left primitives before this chunk within the task; right = position - left
    uint32_t nl = leftScan[blockIdx.x] - leftScan[ch.first[ti]];
    uint32_t nr = c0 - nl;
    for (uint32_t base = c0; base < c1; base += kLB) {
        uint32_t i = base + threadIdx.x;
        bool valid = i < c1, left = false;
        uint32_t p = 0;
        if (valid) {
            p = idxIn[t.begin + i];
            left = goesLeft(s, t, plo[p], phi[p], t.begin + i);
        }
        uint32_t totL, totR;
        uint32_t ol = blockPrefix(left, sw, totL);
        uint32_t orr = blockPrefix(valid && !left, sw, totR);
        if (left) idxOut[t.begin + nl + ol] = p;
        else if (valid) idxOut[t.begin + s.nLeft + nr + orr] = p;
        nl += totL;
        nr += totR;
    }
}

inline unsigned nblk(uint32_t n, unsigned b) { return (n + b - 1) / b; }

#define APH_CHECK(x)                                  \
    do {                                              \
        hipError_t e_ = (x);                          \
        if (e_ != hipSuccess) return e_;              \
    } while (0)

// Device scratch of one build, freed on scope exit.
struct Scratch {
    std::vector<void*> p;
    template <class T>
    hipError_t get(T** out, size_t count)
    {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, count * sizeof(T) + 16);
        if (e == hipSuccess) p.push_back(q);
        *out = (T*)q;
        return e;
    }
    ~Scratch()
    {
        for (void* q : p) hipFree(q);
    }
};

}  // namespace

hipError_t sahBuild(const float4* primLo, const float4* primHi, uint32_t n, uint32_t maxLeaf, bool leafOriginal,
                    const uint32_t* leafRemap, hipStream_t st, BvhNode* nodes, uint32_t* sortedOut,
                    BuildInfo* info)
{
    info->depth = 0;
    info->nodes = 0;
    if (n < 2) return lbvhBuild(primLo, primHi, n, leafOriginal, leafRemap, st, nodes, sortedOut, info);
    if (maxLeaf < 1) maxLeaf = 1;
    if (maxLeaf > kMaxLeafPrims) maxLeaf = kMaxLeafPrims;
    if (leafOriginal) maxLeaf = 1;
    Scratch sc;
    uint32_t *idxA, *idxB, *b12, *counts, *countsNext;
    Task *smallA, *smallB, *largeA, *largeB;
    const uint32_t maxLarge = n / (kSmallMax + 1) + 1;
    const uint32_t maxChunks = n / kChunk + maxLarge + 1;
    uint32_t *gbins, *chunkCnt, *chunkFirst, *chunkTask, *leftCnt, *leftScan;
    Split* splits;
    APH_CHECK(sc.get(&idxA, n));
    APH_CHECK(sc.get(&idxB, n));
    APH_CHECK(sc.get(&b12, 12));
    APH_CHECK(sc.get(&counts, 4));
    APH_CHECK(sc.get(&countsNext, 4));
    APH_CHECK(sc.get(&smallA, n));
    APH_CHECK(sc.get(&smallB, n));
    APH_CHECK(sc.get(&largeA, maxLarge));
    APH_CHECK(sc.get(&largeB, maxLarge));
    APH_CHECK(sc.get(&gbins, (size_t)maxLarge * kTaskBinWords));
    APH_CHECK(sc.get(&chunkCnt, maxLarge));
    APH_CHECK(sc.get(&chunkFirst, maxLarge + 1));
    APH_CHECK(sc.get(&chunkTask, maxChunks));
    APH_CHECK(sc.get(&leftCnt, maxChunks));
    APH_CHECK(sc.get(&leftScan, maxChunks));
    APH_CHECK(sc.get(&splits, maxLarge));
    size_t scanBytes = 0, scanBytes2 = 0;
    APH_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, scanBytes, chunkCnt, chunkFirst, (int)maxLarge, st));
    APH_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, scanBytes2, leftCnt, leftScan, (int)maxChunks, st));
    void* scanTmp;
    APH_CHECK(sc.get((uint8_t**)&scanTmp, std::max(scanBytes, scanBytes2)));
    size_t scanCap = std::max(scanBytes, scanBytes2);

    uint32_t init12[12];
    for (int k = 0; k < 12; k++) init12[k] = (k % 6) < 3 ? 0xFFFFFFFFu : 0u;
    APH_CHECK(hipMemcpyAsync(b12, init12, sizeof(init12), hipMemcpyHostToDevice, st));
    kRootBounds<<<std::min(nblk(n, 256), 512u), 256, 0, st>>>(primLo, primHi, n, b12);
    kIota<<<nblk(n, 256), 256, 0, st>>>(idxA, n);
    kRootTask<<<1, 1, 0, st>>>(b12, n, smallA, largeA, counts);
    APH_CHECK(hipGetLastError());
    uint32_t hc[4];
    APH_CHECK(hipMemcpyAsync(hc, counts, sizeof(hc), hipMemcpyDeviceToHost, st));
    APH_CHECK(hipStreamSynchronize(st));
    uint32_t nSmall = hc[0], nLarge = hc[1], nodeCount = 0;
    uint32_t levels = 0;
    const uint32_t* remap = leafRemap;
    while (nSmall + nLarge > 0) {
        ++levels;
        const int halves = levels > kForceLevel ? 1 : 0;
        if (levels > kForceLevel + 40) return hipErrorUnknown;  // halving ends within log2(n) levels
        // next level's lists start empty; the node allocator carries over
        uint32_t zero4[4] = {0u, 0u, nodeCount, 0u};
        APH_CHECK(hipMemcpyAsync(countsNext, zero4, sizeof(zero4), hipMemcpyHostToDevice, st));
        Lists out{smallB, largeB, countsNext};
        if (nLarge) {
            kChunkCount<<<nblk(nLarge, 256), 256, 0, st>>>(largeA, nLarge, chunkCnt);
            APH_CHECK(hipcub::DeviceScan::ExclusiveSum(scanTmp, scanCap, chunkCnt, chunkFirst, (int)nLarge, st));
            kChunkList<<<nblk(nLarge, 256), 256, 0, st>>>(largeA, nLarge, chunkFirst, chunkTask);
            uint32_t lastFirst = 0, lastCnt = 0;
            APH_CHECK(hipMemcpyAsync(&lastFirst, chunkFirst + nLarge - 1, 4, hipMemcpyDeviceToHost, st));
            APH_CHECK(hipMemcpyAsync(&lastCnt, chunkCnt + nLarge - 1, 4, hipMemcpyDeviceToHost, st));
            APH_CHECK(hipStreamSynchronize(st));
            uint32_t nChunks = lastFirst + lastCnt;
            if (nChunks > maxChunks) return hipErrorUnknown;
            Chunks ch{chunkTask, chunkFirst};
            kBinsInit<<<nblk(nLarge * kTaskBinWords, 256), 256, 0, st>>>(gbins, nLarge);
            kBinLarge<<<nChunks, kLB, 0, st>>>(largeA, ch, primLo, primHi, idxA, gbins);
            kSplitLarge<<<nLarge, 64, 0, st>>>(largeA, gbins, splits, nodes, out, halves);
            kPartCount<<<nChunks, kLB, 0, st>>>(largeA, splits, ch, primLo, primHi, idxA, leftCnt);
            APH_CHECK(hipcub::DeviceScan::ExclusiveSum(scanTmp, scanCap, leftCnt, leftScan, (int)nChunks, st));
            kPartScatter<<<nChunks, kLB, 0, st>>>(largeA, splits, ch, primLo, primHi, idxA, leftScan, idxB);
            APH_CHECK(hipGetLastError());
        }
        if (nSmall) {
            kSmall<<<nSmall, 64, 0, st>>>(smallA, primLo, primHi, idxA, idxB, sortedOut, nodes, out, maxLeaf,
                                          leafOriginal ? 1 : 0, remap, halves);
            APH_CHECK(hipGetLastError());
        }
        APH_CHECK(hipMemcpyAsync(hc, countsNext, sizeof(hc), hipMemcpyDeviceToHost, st));
        APH_CHECK(hipStreamSynchronize(st));
        nSmall = hc[0];
        nLarge = hc[1];
        nodeCount = hc[2];
        if (nodeCount > n - 1) return hipErrorUnknown;
        std::swap(idxA, idxB);
        std::swap(smallA, smallB);
        std::swap(largeA, largeB);
    }
    float rb[12];
    uint32_t b12h[12];
    APH_CHECK(hipMemcpyAsync(b12h, b12, sizeof(b12h), hipMemcpyDeviceToHost, st));
    APH_CHECK(hipStreamSynchronize(st));
    for (int k = 0; k < 12; k++) {
        uint32_t u = b12h[k];
        uint32_t b = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
        memcpy(&rb[k], &b, 4);
    }
    for (int k = 0; k < 3; k++) {
        info->rootLo[k] = rb[k];
        info->rootHi[k] = rb[3 + k];
    }
    info->depth = levels;
    info->nodes = nodeCount;
    return hipSuccess;
}

}  // namespace aph
