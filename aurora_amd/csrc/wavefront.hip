// Wavefront path tracing for gfx950.
//
// The reference runs one megakernel thread per pixel-sample (RayGenShader,
// Shaders/MainEntryPoints.slang:70-334, dispatched W x H per sample by
// HGIRenderer.cpp:551).  Here the same path is split into stages over SoA
// queues shared by every request and sample of a chunk:
//
//   gen      camera rays for all (request, sample, pixel) paths of the chunk
//   extend   closest-hit traversal of the live ray queue          (lean, high occupancy)
//   shade    material, opacity coin, lights, environment MIS, next ray, roulette;
//            appends shadow rays with wave-ballot compaction, marks continuing paths
//   shadow   any-hit traversal of the shadow queue
//   resolve  adds shadow-weighted lighting in the reference's order, routes it
//            to the direct/diffuse/glossy outputs, applies the terminal bounce,
//            and moves the continuing rays to the next queue in queue order
//   accum    per pixel, folds the chunk's samples into the target in sample order
//
// Arithmetic per path is the megakernel's (pathtrace.h helpers, same operation
// order), so a chunk's result does not depend on what else shares the queues.
#include "wavefront.h"

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>

namespace aph {

namespace {

constexpr int kWB = 128;
static_assert(kWB == kStackLanes, "the traversal stack stride is the block size");
constexpr int kSlots = 6;           // shadow slots per path: 0..3 distant lights, 4 MIS material, 5 MIS light
constexpr int kSlotMisMat = 4, kSlotMisLight = 5;

// Event of the last traced segment.
constexpr uint32_t kEvMiss = 0, kEvHit = 1, kEvSkip = 2;

// Packed per-path flags.
struct Flags {
    uint32_t v;
    APH_HD uint32_t get(int sh, int bits) const { return (v >> sh) & ((1u << bits) - 1u); }
    APH_HD void set(int sh, int bits, uint32_t x) { v = (v & ~(((1u << bits) - 1u) << sh)) | ((x & ((1u << bits) - 1u)) << sh); }
    APH_HD uint32_t rayType() const { return get(0, 4); }
    APH_HD void rayType(uint32_t x) { set(0, 4, x); }
    APH_HD uint32_t nextRayType() const { return get(4, 4); }
    APH_HD void nextRayType(uint32_t x) { set(4, 4, x); }
    APH_HD uint32_t misLobe() const { return get(8, 4); }
    APH_HD void misLobe(uint32_t x) { set(8, 4, x); }
    APH_HD uint32_t depth() const { return get(12, 4); }
    APH_HD void depth(uint32_t x) { set(12, 4, x); }
    APH_HD uint32_t event() const { return get(16, 2); }
    APH_HD void event(uint32_t x) { set(16, 2, x); }
    APH_HD bool primary() const { return get(18, 1); }
    APH_HD void primary(bool x) { set(18, 1, x); }
    APH_HD bool ended() const { return get(19, 1); }
    APH_HD void ended(bool x) { set(19, 1, x); }
    APH_HD bool terminal() const { return get(20, 1); }
    APH_HD void terminal(bool x) { set(20, 1, x); }
    APH_HD float alpha() const { return get(21, 1) ? 1.0f : 0.0f; }
    APH_HD void alpha(float a) { set(21, 1, a != 0.0f); }
    APH_HD bool inside() const { return get(22, 1); }
    APH_HD void inside(bool x) { set(22, 1, x); }
};

struct Paths {
    uint32_t* pix;
    uint32_t* req;         // request index of the path
    uint2* rng;
    float4* contrib;
    uint32_t* flags;
    float4* outDirect;
    float4* outDiffuse;
    float4* outGlossy;
    float4* lastIndirect;
    float4* pendDirect;    // emission (hit) or background (miss) of the last segment
    float4* pendContrib;   // contribution that scales this segment's direct / indirect
    float4* nextDir;       // direction of the continuing / terminal ray
    float4* shColor;       // [C * kSlots]
    uint8_t* shMask;
    uint8_t* vis;          // [C * kSlots]
    float4* sampleOut;
};

struct Queues {
    float4* rayA[2];       // org.xyz, ray tag (request index; camera bit)
    float4* rayB[2];       // dir.xyz, path id bits
    float4* hitA;          // t, inst, prim, u
    float* hitV;
    float4* shA;           // org.xyz, request index bits (traced from kRayTMin)
    float4* shB;           // dir.xyz, (path * kSlots + slot) bits
    uint32_t* counts;      // [0],[1] ray queues, [2] shadow queue, [3]/[4] extend/shadow fetch cursors
    uint8_t* cont;         // per entry of the current ray queue: 1 if its path continues
    uint32_t* tileCnt;     // continuing entries per kTile entries of the current queue
    uint32_t* tileOff;     // their exclusive prefix sum: where each tile's survivors go
};

// The next ray queue is written in the order of the current one (a stable
// compaction per bounce), so it stays sorted by path id and the path state
// gathered by path id is read in nearly contiguous runs.
constexpr uint32_t kTile = 1024;

struct SegInfo {
    uint32_t req, sample, pathOffset, count;
};

// Sizes of the chunk in flight, written with its tables: the kernels read them
// from device memory so that one captured launch graph serves every chunk.
struct ChunkHdr {
    uint32_t total, nseg, ncreq, totalPix;
};

struct ChunkReq {
    uint32_t req, firstSeg, nSeg, pixelBase;
};

struct DevReq {
    RenderScene rs;
    aph_frame fp;
    uint32_t W, H, seedOffset, pad;
    float4* accum;
};

// The w word of a ray queue entry: the path's request index, with bit 31 set
// for camera rays, which are traced from t = 0 (every other ray from kRayTMin).
// Carrying the request in the entry lets a kernel fetch its request record
// alongside the ray instead of after a load of path state.
constexpr uint32_t kCameraTag = 0x80000000u;
__device__ __forceinline__ float rayTag(uint32_t req, bool camera)
{
    return __uint_as_float(req | (camera ? kCameraTag : 0u));
}
__device__ __forceinline__ uint32_t tagReq(float w) { return __float_as_uint(w) & ~kCameraTag; }
__device__ __forceinline__ float tagTmin(float w) { return (__float_as_uint(w) & kCameraTag) ? 0.0f : kRayTMin; }

__device__ __forceinline__ uint32_t findSeg(const SegInfo* segs, uint32_t nseg, uint32_t p)
{
    uint32_t lo = 0, hi = nseg;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (segs[mid].pathOffset <= p) lo = mid;
        else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ uint32_t waveAppend(bool keep, uint32_t* counter)
{
    uint64_t active = __ballot(1);
    uint64_t mask = __ballot(keep);
    int lane = __lane_id();
    int leader = __ffsll((unsigned long long)active) - 1;
    uint32_t base = 0;
    if (lane == leader && mask) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, leader, 64);
    return base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
}

// Appends n (0 to 3) entries per active lane with one atomic per wave; returns the
// lane's first position.
__device__ __forceinline__ uint32_t waveAppend3(uint32_t n, uint32_t* counter)
{
    const uint64_t active = __ballot(1);
    const uint64_t b0 = __ballot(n & 1u), b1 = __ballot(n & 2u);
    const int lane = __lane_id();
    const uint64_t below = (1ull << lane) - 1ull;
    const int leader = __ffsll((unsigned long long)active) - 1;
    const uint32_t total = (uint32_t)(__popcll(b0) + 2 * __popcll(b1));
    uint32_t base = 0;
    if (lane == leader && total) base = atomicAdd(counter, total);
    base = __shfl(base, leader, 64);
    return base + (uint32_t)(__popcll(b0 & below) + 2 * __popcll(b1 & below));
}

__device__ __forceinline__ uint32_t waveSum(uint32_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ void addCounter(unsigned long long* c, uint32_t v)
{
    uint32_t s = waveSum(v);
    if ((threadIdx.x & 63) == 0 && s) atomicAdd(c, (unsigned long long)s);
}

__device__ __forceinline__ f3 xyz(float4 v) { return f3{v.x, v.y, v.z}; }
__device__ __forceinline__ float4 f4(f3 v, float w) { return make_float4(v.x, v.y, v.z, w); }

__global__ void kReset(uint32_t* counts, const ChunkHdr* hdr)
{
    if (threadIdx.x == 0) {
        counts[0] = hdr->total;
        counts[1] = 0;
        counts[2] = 0;
        counts[3] = 0;
        counts[4] = 0;
    }
}

__global__ void kResetNext(uint32_t* counts, int next)
{
    if (threadIdx.x == 0) {
        counts[next] = 0;
        counts[2] = 0;
        counts[3] = 0;
        counts[4] = 0;
    }
}

// ---- gen: RayGenShader prologue (MainEntryPoints.slang:83-103) ----
__global__ __launch_bounds__(kWB) void kGen(const DevReq* reqs, const SegInfo* segs, const ChunkHdr* hdr, Paths P,
                                           Queues Q, DevCounters* ctr)
{
    const uint32_t nseg = hdr->nseg, total = hdr->total;
    uint32_t started = 0;
    for (uint32_t p = blockIdx.x * kWB + threadIdx.x; p < total; p += gridDim.x * kWB) {
        uint32_t s = findSeg(segs, nseg, p);
        SegInfo sg = segs[s];
        const DevReq& rq = reqs[sg.req];
        uint32_t pix = p - sg.pathOffset;
        uint32_t px = pix % rq.W, py = pix / rq.W;
        Rng rng = rngInit(sg.sample + rq.seedOffset, rq.W, px, py);
        f3 org, dir;
        cameraRay(rq.fp, px, py, rq.W, rq.H, rng, org, dir);
        P.pix[p] = pix;
        P.req[p] = sg.req;
        P.rng[p] = make_uint2(rng.x, rng.y);
        P.contrib[p] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
        Flags f{0};
        f.rayType(kCameraRay);
        f.nextRayType(kCameraRay);
        f.misLobe(kNoSecondaryRay);
        f.primary(true);
        P.flags[p] = f.v;
        // outDirect and lastIndirect are first written by the first resolve
        float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        P.outDiffuse[p] = z;
        P.outGlossy[p] = z;
        Q.rayA[0][p] = f4(org, rayTag(sg.req, true));
        Q.rayB[0][p] = f4(dir, __uint_as_float(p));
        started++;
    }
    addCounter(&ctr->paths, started);
}

// ---- persistent traversal of a ray queue ----
constexpr uint32_t kNoRay = 0xFFFFFFFFu;

// Each lane keeps one ray in flight and advances it one traversal round per
// iteration.  When at least `refill` lanes of the wave are idle (or all are),
// they start that many queue entries - from the wave's claimed reserve, topped
// up with one atomic on the fetch cursor when it runs short - so lanes do not
// wait for the slowest ray of their wave.  A refill runs the
// ray set-up on a few lanes while the rest wait, so refilling on every idle
// lane costs more than it saves; the threshold is a launch parameter (32 by
// default, searched by aph_wavefront_autotune: it was 16 until measured; 32 gave
// +14 % on the 1080p headline frame, 64 +24 % on 64 small cameras).  The loop
// is wave-uniform: it ends once the cursor has passed n and no lane has a ray.
// load(i, org, dir, tmin) returns entry i's scene; store(i, state) writes its result.
constexpr uint32_t kReserve = 128;  // claimed, not yet started queue entries a wave keeps (two waves' worth)

// Queue entry k of a ray queue (A, B records) copied by the DMA path of
// global_load_lds_dwordx4: lane l of the wave writes its 16 bytes to dst + 16 l.
__device__ __forceinline__ void prefetchToLds(const float4* src, float4* dst)
{
    __builtin_amdgcn_global_load_lds((const void*)src, (__attribute__((address_space(3))) void*)dst, 16, 0, 0);
}

template <bool ANY, bool STATS, class Decode, class Store>
__device__ __forceinline__ uint32_t traceQueue(uint32_t n, uint32_t* cursor, uint32_t refill, uint32_t leafWait,
                                               const StackMem& mem,
                                               TraceStats& ts, const float4* qa, const float4* qb, float4* ringA,
                                               float4* ringB, Decode decode, Store store)
{
    const int lane = __lane_id();
    const uint64_t below = (1ull << lane) - 1ull;
    const uint32_t width = (uint32_t)__popcll(__ballot(1));
    uint32_t ray = kNoRay, done = 0;
    const TraceScene* sc = nullptr;
    TraceState s;
    // Queue entries the wave has claimed but not started (wave-uniform): a claim
    // takes what the refill lacks plus kReserve entries, so most refills start
    // rays without waiting for an atomic (reserve 64: -8 % on 64 small cameras,
    // 256: -1 % on the 1080p frame against 128).  The reserve's rays are copied
    // into the wave's LDS ring (slot = entry - poolStart) while traversal goes
    // on, so a refill from it reads LDS instead of waiting on global memory.
    uint32_t poolBase = 0, poolLeft = 0, poolStart = 0;
    bool exhausted = false;    // the shared cursor has passed n
    bool ringPending = false;  // a copy into the ring may still be in flight
    while (true) {
        const uint64_t idle = __ballot(ray == kNoRay);
        const uint32_t nIdle = (uint32_t)__popcll(idle);
        if ((poolLeft > 0 || !exhausted) && (nIdle >= refill || nIdle == width)) {
            uint32_t newBase = 0, fresh = 0;
            if (poolLeft < nIdle && !exhausted) {
                fresh = nIdle - poolLeft + kReserve;
                const int leader = __ffsll((unsigned long long)idle) - 1;
                if (lane == leader) newBase = atomicAdd(cursor, fresh);
                newBase = __shfl(newBase, leader, 64);
                if (newBase + fresh >= n) exhausted = true;
            }
            if (poolLeft > 0 && ringPending) {
                __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): the ring's copies have landed
                ringPending = false;
            }
            // idle lane of rank r: the reserve's entry r (from the ring), then the fresh claim's
            if (ray == kNoRay) {
                const uint32_t r = (uint32_t)__popcll(idle & below);
                const bool fromPool = r < poolLeft;
                const uint32_t i = fromPool ? poolBase + r : newBase + (r - poolLeft);
                if ((fromPool || r - poolLeft < fresh) && i < n) {
                    float4 a, b;
                    if (fromPool) {
                        a = ringA[i - poolStart];
                        b = ringB[i - poolStart];
                    } else {
                        a = qa[i];
                        b = qb[i];
                    }
                    f3 org, dir;
                    float tmin;
                    bool transparent = false;
                    sc = decode(a, b, org, dir, tmin, transparent);
                    ray = i;
                    const bool live = traceBegin(s, *sc, org, dir, tmin, INFINITY);
                    s.transparent = transparent;
                    if (!live) {
                        store(i, s);
                        ray = kNoRay;
                        done++;
                    }
                }
            }
            if (poolLeft >= nIdle) {
                poolBase += nIdle;
                poolLeft -= nIdle;
            } else {
                const uint32_t used = nIdle - poolLeft;  // taken from the fresh claim
                poolBase = newBase + used;
                poolLeft = fresh >= used ? fresh - used : 0u;
            }
            poolLeft = poolBase >= n ? 0u : min(poolLeft, n - poolBase);
            if (fresh > 0 && poolLeft > 0) {
                // a new reserve (at most kReserve entries from poolBase): copy its rays into the ring
                poolStart = poolBase;
#pragma unroll
                for (uint32_t h = 0; h < kReserve / 64; h++) {
                    const uint32_t k = h * 64 + (uint32_t)lane;
                    if (k < poolLeft) {
                        prefetchToLds(qa + poolStart + k, ringA + h * 64);
                        prefetchToLds(qb + poolStart + k, ringB + h * 64);
                    }
                }
                ringPending = true;
            }
        }
        if (ray != kNoRay && traceRound<ANY, STATS>(s, *sc, mem, ts, leafWait)) {
            store(ray, s);
            ray = kNoRay;
            done++;
        }
        if (exhausted && poolLeft == 0 && !__any(ray != kNoRay)) break;
    }
    return done;
}

// ---- extend: closest hit of the live queue ----
template <bool STATS>
__global__ __launch_bounds__(kWB) void kExtend(const DevReq* reqs, const SegInfo* segs, Paths P, Queues Q, int cur,
                                              uint32_t refill, uint32_t leafWait, DevCounters* ctr)
{
    __shared__ uint32_t sStack[kLdsStack * kWB];
    __shared__ float4 sRing[kWB / 64][2][kReserve];
    uint32_t spill[kSpillStack];
    const StackMem mem{sStack + threadIdx.x, spill};
    TraceStats ts = {};
    const uint32_t n = Q.counts[cur];
    uint32_t found = 0;
    auto decode = [&](float4 a, float4 b, f3& org, f3& dir, float& tmin, bool& transparent) -> const TraceScene* {
        org = xyz(a);
        dir = xyz(b);
        tmin = tagTmin(a.w);
        transparent = false;
        return &reqs[tagReq(a.w)].rs.sc;
    };
    auto store = [&](uint32_t i, const TraceState& s) {
        if (STATS) found += s.found ? 1u : 0u;
        if (s.found) {
            Q.hitA[i] = make_float4((float)s.best.t, __uint_as_float(s.best.inst), __uint_as_float(s.best.prim),
                                    s.best.u);
            Q.hitV[i] = s.best.v;
        } else {
            Q.hitA[i] = make_float4(INFINITY, __uint_as_float(0xFFFFFFFFu), __uint_as_float(0xFFFFFFFFu), 0.0f);
            Q.hitV[i] = 0.0f;
        }
    };
    float4(*ring)[kReserve] = sRing[threadIdx.x >> 6];
    uint32_t done = traceQueue<false, STATS>(n, &Q.counts[3], refill, leafWait, mem, ts, Q.rayA[cur], Q.rayB[cur], ring[0],
                                             ring[1], decode, store);
    addCounter(&ctr->ext, done);
    if (STATS) {
        addCounter(&ctr->nodes, ts.nodes);
        addCounter(&ctr->tris, ts.tris);
        addCounter(&ctr->insts, ts.insts);
        addCounter(&ctr->hits, found);
    }
    addCounter(&ctr->overflow, ts.overflow);
}

// ---- ground plane of primary rays (evaluateGroundPlane, GroundPlane.slang:232-290) ----
// Runs between extend and shade when a request of the call has a ground plane.  For a
// primary ray that meets the plane in front of its hit, it draws the plane's coin and
// samples, traces the matte shadow ray (opaque, per lane), and tints the path's
// contribution.  A reflection ends the segment at the plane: its next ray replaces the
// queue entry and the hit record is marked kGroundInst, which kShade resolves as an
// empty segment (no emission, lights or roulette), as the reference's early return.
constexpr uint32_t kGroundInst = 0xFFFFFFFEu;
template <bool STATS>
__global__ __launch_bounds__(kWB) void kGround(const DevReq* reqs, Paths P, Queues Q, int cur, DevCounters* ctr)
{
    __shared__ uint32_t sStack[kLdsStack * kWB];
    uint32_t spill[kSpillStack];
    const StackMem mem{sStack + threadIdx.x, spill};
    TraceStats ts = {};
    const uint32_t n = Q.counts[cur];
    uint32_t shadows = 0;
    for (uint32_t i = blockIdx.x * kWB + threadIdx.x; i < n; i += gridDim.x * kWB) {
        const float4 rb = Q.rayB[cur][i];
        const uint32_t p = __float_as_uint(rb.w);
        Flags f{P.flags[p]};
        if (!f.primary()) continue;
        const float4 ra = Q.rayA[cur][i];
        const uint32_t ri = tagReq(ra.w);
        const DevReq& rq = reqs[ri];
        const DevGround& g = rq.rs.ground;
        f3 pos;
        float dist;
        if (!g.enabled || !groundPlaneHit(g, xyz(ra), xyz(rb), pos, dist) || !(dist < Q.hitA[i].x)) continue;
        const uint2 rs = P.rng[p];
        Rng rng{rs.x, rs.y};
        f3 color, ndir, sdir;
        bool shadow;
        const int type = groundPlaneSetup(g, rq.rs.env, xyz(rb), rng, color, ndir, shadow, sdir);
        if (shadow) {
            color = color * groundShadowTint(g, traceVisibility<STATS>(rq.rs.sc, pos, sdir, kRayTMin, false, mem, ts));
            shadows++;
        }
        P.contrib[p] = f4(xyz(P.contrib[p]) * color, 0.0f);
        P.rng[p] = make_uint2(rng.x, rng.y);
        if (type == kGroundPlaneReflectionRay) {
            f.primary(false);
            f.nextRayType((uint32_t)kGroundPlaneReflectionRay);
            P.flags[p] = f.v;
            Q.rayA[cur][i] = f4(pos, rayTag(ri, false));
            Q.rayB[cur][i] = f4(ndir, __uint_as_float(p));
            Q.hitA[i] = make_float4(INFINITY, __uint_as_float(kGroundInst), __uint_as_float(0xFFFFFFFFu), 0.0f);
        }
    }
    addCounter(&ctr->shadow, shadows);
    if (STATS) {
        addCounter(&ctr->nodes, ts.nodes);
        addCounter(&ctr->tris, ts.tris);
        addCounter(&ctr->insts, ts.insts);
        addCounter(&ctr->snodes, ts.nodes);
        addCounter(&ctr->stris, ts.tris);
        addCounter(&ctr->sinsts, ts.insts);
    }
    addCounter(&ctr->overflow, ts.overflow);
}

__device__ __forceinline__ void emitShadow(const Queues& Q, bool emit, f3 org, f3 dir, uint32_t slotId, uint32_t req)
{
    uint32_t at = waveAppend(emit, &Q.counts[2]);
    if (emit) {
        Q.shA[at] = f4(org, __uint_as_float(req));
        Q.shB[at] = f4(dir, __uint_as_float(slotId));
    }
}

// ---- shade: closest-hit / miss programs without the shadow traces ----

// Miss program of queue entry i (the path ends; ShadeFunctions.slang via
// MainEntryPoints.slang:385-405): background for primary, transmission and
// ground-plane reflection rays.  Contribution and RNG state are unchanged.
__device__ __forceinline__ void shadeMiss(const DevReq* reqs, const SegInfo* segs, const Paths& P, const Queues& Q,
                                          int cur, uint32_t i)
{
    const uint32_t p = __float_as_uint(Q.rayB[cur][i].w);
    Flags f{P.flags[p]};
    const bool bg = f.primary() || f.rayType() == (uint32_t)kTransmissionLobe ||
                    f.rayType() == (uint32_t)kGroundPlaneReflectionRay;
    f3 d = splat(0.0f);
    if (bg) {
        const DevReq& rq = reqs[tagReq(Q.rayA[cur][i].w)];
        const uint32_t pix = P.pix[p];
        d = evalEnvironment(rq.rs.env, xyz(Q.rayB[cur][i]), true, pix % rq.W, pix / rq.W, rq.W, rq.H) *
            xyz(P.contrib[p]);
    }
    P.pendDirect[p] = f4(d, 0.0f);
    f.event(kEvMiss);
    f.ended(true);
    f.terminal(false);
    P.shMask[p] = 0;
    P.flags[p] = f.v;
    Q.cont[i] = 0;
}

// A segment that ended at the ground plane (kGround): nothing is shaded; the path
// continues with the reflection ray kGround put in entry i, unless the depth runs out.
__device__ __forceinline__ void shadeGround(const DevReq* reqs, const Paths& P, const Queues& Q, int cur, uint32_t i)
{
    const float4 rb = Q.rayB[cur][i];
    const uint32_t p = __float_as_uint(rb.w);
    Flags f{P.flags[p]};
    const DevReq& rq = reqs[tagReq(Q.rayA[cur][i].w)];
    P.pendDirect[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    P.pendContrib[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    P.shMask[p] = 0;
    f.event(kEvHit);
    f.ended(false);
    bool cont = true;
    f.terminal(false);
    if ((int)f.depth() + 1 >= rq.fp.trace_depth) {
        f.terminal(true);
        cont = false;
        P.nextDir[p] = f4(xyz(rb), 0.0f);
    }
    P.flags[p] = f.v;
    Q.cont[i] = cont ? 1 : 0;
}

// Closest-hit program of queue entry i (live lanes).  A continuing path's next
// ray replaces entry i of the current queue (the path id stays in rayB.w), and
// kResolve moves it to the next queue in queue order.
__device__ __forceinline__ void shadeHit(const DevReq* reqs, const SegInfo* segs, const Paths& P, const Queues& Q,
                                         int cur, bool live, uint32_t i)
{
    bool contPath = false;
    f3 nOrg = splat(0.0f), nDir = splat(0.0f);
    uint32_t p = 0, ri = 0;
    if (live) {
        float4 rb = Q.rayB[cur][i];
        p = __float_as_uint(rb.w);
        ri = tagReq(Q.rayA[cur][i].w);
        const DevReq& rq = reqs[ri];
        const aph_frame& fp = rq.fp;
        const DevEnv& env = rq.rs.env;
        uint32_t pix = P.pix[p], px = pix % rq.W, py = pix / rq.W;
        Flags f{P.flags[p]};
        f3 dir = xyz(rb);
        f3 contrib = xyz(P.contrib[p]);
        uint2 rs = P.rng[p];
        Rng rng{rs.x, rs.y};
        float4 ha = Q.hitA[i];
        uint32_t inst = __float_as_uint(ha.y);
        uint32_t depth = f.depth();
        uint8_t mask = 0;
        // A shadow request is queued where it is made (the appends work on
        // whatever lanes are active), so no slot state stays live.
        auto shadow = [&](int slot, f3 org, f3 sdir, f3 color) {
            P.shColor[(size_t)p * kSlots + slot] = f4(color, 0.0f);
            mask |= (uint8_t)(1u << slot);
            emitShadow(Q, true, org, sdir, p * kSlots + (uint32_t)slot, ri);
        };
        {
            f3 V = -dir;
            const InstShade& ins = rq.rs.ishade[inst];
            Shading shd;
            const uint32_t prim = __float_as_uint(ha.z);
            const float hv = Q.hitV[i];
            computeShading(ins, prim, ha.w, hv, shd);
            HitUV huv(ins, prim, ha.w, hv);
            float Pt;
            f3 opacity;
            const uint32_t mi = selectLayer(rq.rs, ins, huv, rng, contrib, Pt, opacity);
            if (mi == kNoMaterial) {
                nOrg = shd.geomPosition;
                nDir = dir;
                f.event(kEvSkip);
                contPath = true;
            } else {
                const DevMaterial& m = rq.rs.mats[mi];
                f3 emissionColor;
                materialTextures(rq.rs, mi, m, huv, shd, emissionColor);
                f3 direct = emissionColor * m.emission * lerp(splat(1.0f), m.coatColor, m.coat);
                f3 so = dot(V, shd.normal) < 0.0f ? shd.geomPosition : shd.position;
                f.primary(false);
                if (fp.light_count > 1) {
                    for (int li = 0; li < fp.light_count; li++) {
                        if (fp.lights[li].color_intensity[3] <= 0.0f) continue;
                        f3 sd;
                        f3 lr = shadeDirectional(fp.lights[li], m, shd, V, rng, sd);
                        if (isBlack(lr)) continue;
                        shadow(li, so, sd, lr);
                    }
                }
                // the single light's shadow ray is queued with the environment's below
                bool emitL0 = false;
                f3 l0Dir = splat(0.0f);
                if (fp.light_count == 1) {
                    const aph_distant_light& l0 = fp.lights[0];
                    float w0 = luminance(f3{l0.color_intensity[0], l0.color_intensity[1], l0.color_intensity[2]}) *
                               l0.color_intensity[3];
                    float selPdf = 1.0f;
                    random2D(rng);
                    if (w0 > 0.0f) selPdf = w0 / w0;
                    if (l0.color_intensity[3] > 0.0f) {
                        f3 dc = shadeDirectional(l0, m, shd, V, rng, l0Dir) / fmaxf(selPdf, 1e-4f);
                        P.shColor[(size_t)p * kSlots] = f4(dc, 0.0f);
                        mask |= 1u;
                        emitL0 = true;
                    }
                }
                // environment MIS (ShadeFunctions.slang:135-259).  Its shadow rays and the
                // single light's are queued together, with one atomic per wave.
                {
                    f2 r = random2D(rng);
                    f3 L;
                    float matPdf;
                    int lobe;
                    f3 fs = sampleMaterial(m, shd, V, r, L, matPdf, lobe);
                    bool emitMat = false, emitLight = false;
                    f3 matDir = splat(0.0f), matColor = splat(0.0f), lightDir = splat(0.0f), lightColor = splat(0.0f);
                    if (lobe != kTransmissionLobe) {
                        f3 Le = evalEnvironment(env, L, false, px, py, rq.W, rq.H);
                        if (!isBlack(Le) && !isBlack(fs)) {
                            float lightPdf = env.hasLightTex ? luminance(Le) / env.lightIntegral : 0.25f * kPiInv;
                            emitMat = true;
                            matDir = L;
                            matColor = Le * misWeight(matPdf, lightPdf) * fs / matPdf;
                        }
                        float lightPdf;
                        Le = sampleEnvironment(env, r, L, lightPdf);
                        int evLobe;
                        fs = evalMaterialPdf(m, shd, V, L, r, matPdf, evLobe);
                        if (!isBlack(Le) && !isBlack(fs)) {
                            emitLight = true;
                            lightDir = L;
                            lightColor = Le * misWeight(lightPdf, matPdf) * fs / lightPdf;
                        }
                    }
                    uint32_t at = waveAppend3((uint32_t)emitL0 + (uint32_t)emitMat + (uint32_t)emitLight, &Q.counts[2]);
                    if (emitL0) {
                        Q.shA[at] = f4(so, __uint_as_float(ri));
                        Q.shB[at] = f4(l0Dir, __uint_as_float(p * kSlots));
                        at++;
                    }
                    if (emitMat) {
                        P.shColor[(size_t)p * kSlots + kSlotMisMat] = f4(matColor, 0.0f);
                        Q.shA[at] = f4(so, __uint_as_float(ri));
                        Q.shB[at] = f4(matDir, __uint_as_float(p * kSlots + (uint32_t)kSlotMisMat));
                        mask |= (uint8_t)(1u << kSlotMisMat);
                        at++;
                    }
                    if (emitLight) {
                        P.shColor[(size_t)p * kSlots + kSlotMisLight] = f4(lightColor, 0.0f);
                        Q.shA[at] = f4(so, __uint_as_float(ri));
                        Q.shB[at] = f4(lightDir, __uint_as_float(p * kSlots + (uint32_t)kSlotMisLight));
                        mask |= (uint8_t)(1u << kSlotMisLight);
                    }
                    if (mask & ((1u << kSlotMisMat) | (1u << kSlotMisLight))) f.misLobe((uint32_t)lobe);
                }
                contrib = contrib * (opacity / (1.0f - Pt));
                f.alpha((f.alpha() == 0.0f && f.rayType() == (uint32_t)kTransmissionLobe) ? 0.0f : 1.0f);
                f3 nextContrib = splat(1.0f);
                uint32_t nextType;
                {
                    f2 r = random2D(rng);
                    f3 L;
                    float pdf;
                    int lobe;
                    f3 fs = sampleMaterial(m, shd, V, r, L, pdf, lobe);
                    if (isBlack(fs)) {
                        nextType = kNoSecondaryRay;
                    } else {
                        bool trans = lobe == kTransmissionLobe;
                        nOrg = (dot(V, shd.normal) < 0.0f || trans) ? shd.geomPosition : shd.position;
                        nDir = L;
                        nextContrib = fs / pdf;
                        nextType = (uint32_t)lobe;
                    }
                }
                bool ended = nextType == (uint32_t)kNoSecondaryRay;
                if (nextType == (uint32_t)kTransmissionLobe) f.inside(!f.inside());
                P.pendDirect[p] = f4(direct, 0.0f);
                P.pendContrib[p] = f4(contrib, 0.0f);
                contrib = contrib * nextContrib;
                if (fp.rr_enabled && (int)depth + 1 >= fp.rr_start_depth) {
                    float surv = clampf(maxc(contrib), 0.05f, 1.0f);
                    if (random2D(rng).x > surv) ended = true;
                    else contrib = contrib / surv;
                }
                f.nextRayType(nextType);
                f.event(kEvHit);
                f.ended(ended);
                contPath = !ended;
            }
        }
        f.terminal(false);
        if (contPath) {
            uint32_t nd = depth + 1;
            if ((int)nd >= fp.trace_depth) {
                f.terminal(true);
                contPath = false;
                P.nextDir[p] = f4(nDir, 0.0f);
            }
        }
        P.shMask[p] = mask;
        P.flags[p] = f.v;
        P.contrib[p] = f4(contrib, 0.0f);
        P.rng[p] = make_uint2(rng.x, rng.y);
        Q.cont[i] = contPath ? 1 : 0;
        if (contPath) {
            Q.rayA[cur][i] = f4(nOrg, rayTag(ri, false));
            Q.rayB[cur][i] = f4(nDir, __uint_as_float(p));
        }
    }
}

// The block walks its grid-stride slices of the queue: misses are shaded where
// they are found, hits are queued in LDS and shaded kWB at a time, so the long
// closest-hit program runs on full waves whatever the miss rate.  Results are
// per path, so the order of shading does not change them.  Four waves per SIMD
// (128 VGPRs, a few spills) hide the gathers of shading data, materials and
// environment texels better than three without spills.
__global__ __launch_bounds__(kWB) __attribute__((amdgpu_waves_per_eu(4, 8))) void kShade(const DevReq* reqs, const SegInfo* segs, Paths P, Queues Q, int cur)
{
    __shared__ uint32_t hitq[2 * kWB];
    __shared__ uint32_t hitCount;
    const uint32_t n = Q.counts[cur];
    const uint32_t stride = gridDim.x * kWB;
    uint32_t next = blockIdx.x * kWB;  // start of the block's next slice (block-uniform)
    uint32_t pending = 0;              // hits queued in LDS (block-uniform)
    if (threadIdx.x == 0) hitCount = 0;
    __syncthreads();
    for (;;) {
        // Refill to at least one block of hits (at most 2 kWB - 1 queued).
        while (pending < kWB && next < n) {
            const uint32_t i = next + threadIdx.x;
            next += stride;
            bool hit = false;
            if (i < n) {
                const uint32_t inst = __float_as_uint(Q.hitA[i].y);
                hit = inst < kGroundInst;
                if (inst == kGroundInst) shadeGround(reqs, P, Q, cur, i);
                else if (!hit) shadeMiss(reqs, segs, P, Q, cur, i);
            }
            const uint64_t m = __ballot(hit);
            const int lane = __lane_id();
            uint32_t base = 0;
            if (lane == 0 && m) base = atomicAdd(&hitCount, (uint32_t)__popcll(m));
            base = __shfl(base, 0, 64);
            if (hit) hitq[base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = i;
            __syncthreads();
            pending = hitCount;
            __syncthreads();
        }
        if (pending == 0) break;
        const uint32_t take = pending < (uint32_t)kWB ? pending : (uint32_t)kWB;
        const uint32_t rest = pending - take;
        const bool live = threadIdx.x < take;
        const uint32_t i = live ? hitq[rest + threadIdx.x] : 0u;
        if (threadIdx.x == 0) hitCount = rest;
        pending = rest;
        __syncthreads();
        shadeHit(reqs, segs, P, Q, cur, live, i);
    }
}

// ---- shadow: any-hit of the shadow queue (traceShadowRay; opaque, or transmitted by non-opaque instances) ----
template <bool STATS>
__global__ __launch_bounds__(kWB) void kShadow(const DevReq* reqs, const SegInfo* segs, Paths P, Queues Q,
                                              uint32_t refill, uint32_t leafWait, DevCounters* ctr)
{
    __shared__ uint32_t sStack[kLdsStack * kWB];
    __shared__ float4 sRing[kWB / 64][2][kReserve];
    uint32_t spill[kSpillStack];
    const StackMem mem{sStack + threadIdx.x, spill};
    TraceStats ts = {};
    const uint32_t n = Q.counts[2];
    auto decode = [&](float4 a, float4 b, f3& org, f3& dir, float& tmin, bool& transparent) -> const TraceScene* {
        org = xyz(a);
        dir = xyz(b);
        tmin = kRayTMin;
        const DevReq& rq = reqs[__float_as_uint(a.w)];
        // gFrameData.isForceOpaqueShadowsEnabled (HitShaderEntryPoints.slang:42)
        transparent = rq.fp.force_opaque_shadows == 0 && rq.rs.sc.transmit != nullptr;
        return &rq.rs.sc;
    };
    // A transmitted shadow ray scales its slot's light colour by the visibility it
    // accumulated (vis stays 1), the reference's lightVisibility * colour.
    auto store = [&](uint32_t i, const TraceState& s) {
        const uint32_t id = __float_as_uint(Q.shB[i].w);
        if (s.transparent && !s.found) {
            const float4 c = P.shColor[id];
            P.shColor[id] = make_float4(s.vis.x * c.x, s.vis.y * c.y, s.vis.z * c.z, c.w);
        }
        P.vis[id] = s.found ? 0 : 1;
    };
    float4(*ring)[kReserve] = sRing[threadIdx.x >> 6];
    uint32_t done = traceQueue<true, STATS>(n, &Q.counts[4], refill, leafWait, mem, ts, Q.shA, Q.shB, ring[0], ring[1], decode,
                                            store);
    addCounter(&ctr->shadow, done);
    if (STATS) {
        addCounter(&ctr->nodes, ts.nodes);
        addCounter(&ctr->tris, ts.tris);
        addCounter(&ctr->insts, ts.insts);
        addCounter(&ctr->snodes, ts.nodes);
        addCounter(&ctr->stris, ts.tris);
        addCounter(&ctr->sinsts, ts.insts);
    }
    addCounter(&ctr->overflow, ts.overflow);
}

__device__ __forceinline__ void finalizePath(const Paths& P, uint32_t p, const aph_frame& fp, Flags f)
{
    f3 fin = adjustRadiance(fp.max_luminance, fp.display_errors != 0,
                            xyz(P.outDirect[p]) + xyz(P.outDiffuse[p]) + xyz(P.outGlossy[p]));
    P.sampleOut[p] = f4(fin, f.alpha());
}

// ---- stable compaction of the continuing rays: tile counts, then their prefix sum ----

// One wave per kTile entries of the current queue (kTile / 64 flags per lane).
constexpr uint32_t kFlagsPerLane = kTile / 64;
static_assert(kTile % 256 == 0, "each lane counts whole 32-bit words of flags");
__global__ __launch_bounds__(kWB) void kCount(Queues Q, int cur)
{
    const uint32_t n = Q.counts[cur];
    const uint32_t nt = (n + kTile - 1) / kTile;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t waves = gridDim.x * (kWB / 64);
    for (uint32_t t = blockIdx.x * (kWB / 64) + (threadIdx.x >> 6); t < nt; t += waves) {
        const uint32_t i0 = t * kTile + lane * kFlagsPerLane;
        uint32_t s = 0;
        if (i0 + kFlagsPerLane <= n) {
            const uint32_t* w = (const uint32_t*)(Q.cont + i0);
            // flags are 0 or 1, so the byte sum of a word fits its top byte
#pragma unroll
            for (uint32_t k = 0; k < kFlagsPerLane / 4; k++) s += (w[k] * 0x01010101u) >> 24;
        } else {
            for (uint32_t k = 0; k < kFlagsPerLane; k++)
                if (i0 + k < n) s += Q.cont[i0 + k];
        }
        s = waveSum(s);
        if (lane == 0) Q.tileCnt[t] = s;
    }
}

// One block of kScanThreads: exclusive prefix sum of the tile counts into
// tileOff, and the total as the next queue's length.
constexpr int kScanThreads = 1024;
__global__ __launch_bounds__(kScanThreads) void kScan(Queues Q, int cur)
{
    __shared__ uint32_t part[kScanThreads];
    const uint32_t n = Q.counts[cur];
    const uint32_t nt = (n + kTile - 1) / kTile;
    const uint32_t per = (nt + kScanThreads - 1) / kScanThreads;
    const uint32_t b = threadIdx.x * per, e = min(nt, b + per);
    uint32_t s = 0;
    for (uint32_t t = b; t < e; t++) s += Q.tileCnt[t];
    part[threadIdx.x] = s;
    __syncthreads();
    // Hillis-Steele inclusive scan over the per-thread sums
    for (int off = 1; off < kScanThreads; off <<= 1) {
        const uint32_t add = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += add;
        __syncthreads();
    }
    uint32_t run = part[threadIdx.x] - s;
    for (uint32_t t = b; t < e; t++) {
        const uint32_t c = Q.tileCnt[t];
        Q.tileOff[t] = run;
        run += c;
    }
    if (threadIdx.x == kScanThreads - 1) Q.counts[cur ^ 1] = part[kScanThreads - 1];
}

// The accumulation part of the RayGenShader loop (MainEntryPoints.slang:124-247)
// for queue entry i, after its shadow rays are traced.
__device__ __forceinline__ void resolvePath(const DevReq* reqs, const Paths& P, const Queues& Q, int cur, uint32_t i)
{
    uint32_t p = __float_as_uint(Q.rayB[cur][i].w);
    const DevReq& rq = reqs[P.req[p]];
    const aph_frame& fp = rq.fp;
    Flags f{P.flags[p]};
    uint32_t depth = f.depth();
    uint32_t ev = f.event();
    if (ev != kEvSkip) {
        f3 direct = xyz(P.pendDirect[p]);
        f3 indirect = splat(0.0f);
        if (ev == kEvHit) {
            uint8_t mask = P.shMask[p];
            const size_t b = (size_t)p * kSlots;
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (mask & (1u << k)) direct = direct + xyz(P.shColor[b + k]) * (P.vis[b + k] ? 1.0f : 0.0f);
            if (mask & ((1u << kSlotMisMat) | (1u << kSlotMisLight))) {
                f3 mis = splat(0.0f);
                if (mask & (1u << kSlotMisMat))
                    mis = mis + xyz(P.shColor[b + kSlotMisMat]) * (P.vis[b + kSlotMisMat] ? 1.0f : 0.0f);
                if (mask & (1u << kSlotMisLight))
                    mis = mis + xyz(P.shColor[b + kSlotMisLight]) * (P.vis[b + kSlotMisLight] ? 1.0f : 0.0f);
                indirect = mis;
            }
            f3 c = xyz(P.pendContrib[p]);
            direct = direct * c;
            indirect = indirect * c;
        }
        if (depth == 0) {
            f.primary(false);
            P.outDirect[p] = f4(direct, 0.0f);
            if (f.misLobe() == (uint32_t)kDiffuseLobe) P.outDiffuse[p] = f4(indirect, 0.0f);
            else P.outGlossy[p] = f4(indirect, 0.0f);
        } else {
            f3 radiance = direct + indirect;
            if (f.rayType() == (uint32_t)kDiffuseLobe) P.outDiffuse[p] = f4(xyz(P.outDiffuse[p]) + radiance, 0.0f);
            else P.outGlossy[p] = f4(xyz(P.outGlossy[p]) + radiance, 0.0f);
        }
        if (ev == kEvHit) f.rayType(f.nextRayType());
        P.lastIndirect[p] = f4(indirect, 0.0f);
    } else {
        // a segment through a fully transparent surface: nothing is added
        if (depth == 0) P.outDirect[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        P.lastIndirect[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    bool done = f.ended();
    if (!done && f.terminal()) {
        // rayDepth == traceDepth: unoccluded environment in the next direction plus the
        // previous segment's indirect, which the reference does not clear (:127-138, :198)
        uint32_t pix = P.pix[p];
        bool bg = f.primary() || f.rayType() == (uint32_t)kTransmissionLobe ||
                  f.rayType() == (uint32_t)kGroundPlaneReflectionRay;
        f3 dterm = evalEnvironment(rq.rs.env, xyz(P.nextDir[p]), bg, pix % rq.W, pix / rq.W, rq.W, rq.H) *
                   xyz(P.contrib[p]);
        f3 radiance = dterm + xyz(P.lastIndirect[p]);
        if (f.rayType() == (uint32_t)kDiffuseLobe) P.outDiffuse[p] = f4(xyz(P.outDiffuse[p]) + radiance, 0.0f);
        else P.outGlossy[p] = f4(xyz(P.outGlossy[p]) + radiance, 0.0f);
        done = true;
    }
    f.depth(depth + 1);
    P.flags[p] = f.v;
    if (done) finalizePath(P, p, fp, f);
}

// ---- resolve: the accumulation part of the RayGenShader loop (MainEntryPoints.slang:124-247) ----
// Each wave walks whole tiles of the queue in order and copies the entries of
// continuing paths (their next rays, written by kShade) to the next queue at
// the tile's offset, so the next queue keeps this queue's order.
__global__ __launch_bounds__(kWB) void kResolve(const DevReq* reqs, const SegInfo* segs, Paths P, Queues Q, int cur)
{
    const uint32_t n = Q.counts[cur];
    const uint32_t nt = (n + kTile - 1) / kTile;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t waves = gridDim.x * (kWB / 64);
    for (uint32_t t = blockIdx.x * (kWB / 64) + (threadIdx.x >> 6); t < nt; t += waves) {
        uint32_t at = Q.tileOff[t];
        for (uint32_t j = 0; j < kTile; j += 64) {
            const uint32_t i = t * kTile + j + lane;
            const bool cont = i < n && Q.cont[i] != 0;
            const uint64_t m = __ballot(cont);
            if (cont) {
                const uint32_t o = at + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
                Q.rayA[cur ^ 1][o] = Q.rayA[cur][i];
                Q.rayB[cur ^ 1][o] = Q.rayB[cur][i];
            }
            at += (uint32_t)__popcll(m);
            if (i < n) resolvePath(reqs, P, Q, cur, i);
        }
    }
}

// ---- accum: running mean in sample order (Accumulation.glsl:22-30) ----
__global__ void kAccum(const DevReq* reqs, const SegInfo* segs, const ChunkReq* creqs, const ChunkHdr* hdr, Paths P)
{
    const uint32_t ncreq = hdr->ncreq, totalPix = hdr->totalPix;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < totalPix; t += gridDim.x * blockDim.x) {
        uint32_t lo = 0, hi = ncreq;
        while (hi - lo > 1) {
            uint32_t mid = (lo + hi) >> 1;
            if (creqs[mid].pixelBase <= t) lo = mid;
            else hi = mid;
        }
        ChunkReq cr = creqs[lo];
        const DevReq& rq = reqs[cr.req];
        uint32_t pix = t - cr.pixelBase;
        float4 acc = rq.accum[pix];
        for (uint32_t s = cr.firstSeg; s < cr.firstSeg + cr.nSeg; s++) {
            SegInfo sg = segs[s];
            float4 r = P.sampleOut[sg.pathOffset + pix];
            if (sg.sample > 0) {
                float tt = 1.0f / (float)(sg.sample + 1);
                float u = 1.0f - tt;
                acc = make_float4(acc.x * u + r.x * tt, acc.y * u + r.y * tt, acc.z * u + r.z * tt,
                                  acc.w * u + r.w * tt);
            } else {
                acc = r;
            }
        }
        rq.accum[pix] = acc;
    }
}

size_t alignUp(size_t x) { return (x + 255) & ~(size_t)255; }

// ---- call table: request records and every chunk's tables, uploaded once per call ----
// Layout (byte offsets in CallHdr): CallHdr | DevReq[nreq] | ChunkDesc[nchunk] |
// SegInfo[all segments] | ChunkReq[all chunk requests].  kChunkNext, the first
// node of every chunk's launch sequence, copies chunk `cursor`'s tables to the
// fixed per-chunk addresses the stages read; kChunkBump advances the cursor.
struct CallHdr {
    uint32_t cursor, nchunk, nreq, pad;
    uint32_t descOff, segOff, creqOff, pad2;
};
struct ChunkDesc {
    ChunkHdr hdr;
    uint32_t segBase, creqBase, pad[2];
};
constexpr size_t kCallHdrBytes = 256;  // DevReq records start here

__global__ void kChunkNext(const uint8_t* tbl, ChunkHdr* ghdr, SegInfo* gseg, ChunkReq* gcr)
{
    const CallHdr* call = (const CallHdr*)tbl;
    const uint32_t c = call->cursor;
    const ChunkDesc d = ((const ChunkDesc*)(tbl + call->descOff))[c];
    const SegInfo* segs = (const SegInfo*)(tbl + call->segOff) + d.segBase;
    const ChunkReq* creqs = (const ChunkReq*)(tbl + call->creqOff) + d.creqBase;
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) *ghdr = d.hdr;
    for (uint32_t i = t; i < d.hdr.nseg; i += stride) gseg[i] = segs[i];
    for (uint32_t i = t; i < d.hdr.ncreq; i += stride) gcr[i] = creqs[i];
}

__global__ void kChunkBump(uint8_t* tbl)
{
    if (threadIdx.x == 0) ((CallHdr*)tbl)->cursor += 1;
}

// Host -> device upload through kernel arguments: the runtime copies a launch's
// arguments when the launch is enqueued, so the host image may be freed as soon
// as the call returns and no host memory stays referenced by queued work.
constexpr uint32_t kUploadWords = 504;  // 2 KB of payload per launch
struct UploadArgs {
    uint32_t n, pad[3];
    uint32_t w[kUploadWords];
};

__global__ void kUpload(uint32_t* dst, UploadArgs a)
{
    for (uint32_t i = threadIdx.x; i < a.n; i += blockDim.x) dst[i] = a.w[i];
}

hipError_t upload(void* dst, const std::vector<uint32_t>& words, hipStream_t stream)
{
    UploadArgs a;
    std::memset(&a, 0, sizeof(a));
    for (size_t k = 0; k < words.size(); k += kUploadWords) {
        a.n = (uint32_t)std::min<size_t>(kUploadWords, words.size() - k);
        std::memcpy(a.w, words.data() + k, a.n * 4);
        kUpload<<<1, 256, 0, stream>>>((uint32_t*)dst + k, a);
    }
    return hipGetLastError();
}

constexpr uint32_t kChunkBlocks = 16;

// Bytes per path of the path state and queues (the Paths and Queues arrays).
constexpr size_t kPerPath = 4 + 4 + 8 + 16 + 4 + 16 * 4 + 16 * 3 + 16 * kSlots + 1 + kSlots + 16 +  // path state
                            16 * 4 + 16 + 4 + 32 * kSlots + 1 + 1;                                      // queues
constexpr size_t kTableBytes = 1u << 20;  // the chunk in flight: header + segment + chunk-request tables
constexpr size_t kHdrBytes = 256;
constexpr uint32_t kMaxSegs = (uint32_t)((kTableBytes / 2 - kHdrBytes) / sizeof(SegInfo));
constexpr uint32_t kMaxCreqs = (uint32_t)(kTableBytes / 2 / sizeof(ChunkReq));
constexpr size_t kFixedBytes = 64 * 1024 + kTableBytes;  // alignment padding, tile tables, counters; chunk tables

}  // namespace

namespace {

// Launch sizes of one chunk: persistent kernels with grid-stride loops, sized by
// the chunk capacity (not the chunk's own path count), so the same launches -
// and the same captured graph - serve every chunk.
struct Launch {
    unsigned gen, trace, shade, light, shadow, accum;
    uint32_t refill;    // idle lanes that trigger a traversal refill
    uint32_t leafWait;  // lanes still looking for a leaf when a traversal wave processes leaves
};

// Everything one chunk runs: its tables moved into place from the call table,
// generation, then per bounce the reset, extension, ground plane (when a request
// has one), shading, shadow and resolve stages, then accumulation.  Timed per stage with HIP events when timing is given.
void enqueueChunk(hipStream_t stream, const Launch& L, uint32_t maxDepth, bool stats, bool ground, uint8_t* tbl,
                  const DevReq* dreq, SegInfo* gseg, ChunkReq* gcr, ChunkHdr* hdr, const Paths& P, const Queues& Q,
                  DevCounters* ctr, std::vector<PendingEvent>* timing)
{
    {
        EventScope ev(timing, "wf_gen", stream);
        kChunkNext<<<kChunkBlocks, 256, 0, stream>>>(tbl, hdr, gseg, gcr);
        kChunkBump<<<1, 64, 0, stream>>>(tbl);
        kReset<<<1, 64, 0, stream>>>(Q.counts, hdr);
        kGen<<<L.gen, kWB, 0, stream>>>(dreq, gseg, hdr, P, Q, ctr);
    }
    int cur = 0;
    for (uint32_t d = 0; d < maxDepth; d++) {
        kResetNext<<<1, 64, 0, stream>>>(Q.counts, cur ^ 1);
        {
            EventScope ev(timing, "wf_extend", stream);
            if (stats) kExtend<true><<<L.trace, kWB, 0, stream>>>(dreq, gseg, P, Q, cur, L.refill, L.leafWait, ctr);
            else kExtend<false><<<L.trace, kWB, 0, stream>>>(dreq, gseg, P, Q, cur, L.refill, L.leafWait, ctr);
        }
        if (ground) {
            EventScope ev(timing, "wf_ground", stream);
            if (stats) kGround<true><<<L.light, kWB, 0, stream>>>(dreq, P, Q, cur, ctr);
            else kGround<false><<<L.light, kWB, 0, stream>>>(dreq, P, Q, cur, ctr);
        }
        {
            EventScope ev(timing, "wf_shade", stream);
            kShade<<<L.shade, kWB, 0, stream>>>(dreq, gseg, P, Q, cur);
        }
        {
            EventScope ev(timing, "wf_shadow", stream);
            if (stats) kShadow<true><<<L.shadow, kWB, 0, stream>>>(dreq, gseg, P, Q, L.refill, L.leafWait, ctr);
            else kShadow<false><<<L.shadow, kWB, 0, stream>>>(dreq, gseg, P, Q, L.refill, L.leafWait, ctr);
        }
        {
            EventScope ev(timing, "wf_resolve", stream);
            kCount<<<L.light, kWB, 0, stream>>>(Q, cur);
            kScan<<<1, kScanThreads, 0, stream>>>(Q, cur);
            kResolve<<<L.light, kWB, 0, stream>>>(dreq, gseg, P, Q, cur);
        }
        cur ^= 1;
    }
    {
        EventScope ev(timing, "wf_accum", stream);
        kAccum<<<L.accum, 256, 0, stream>>>(dreq, gseg, gcr, hdr, P);
    }
}

// Segments of a call: (request, sample) pairs in request-major, sample-ascending
// order, cut into chunks of at most `cap` paths and the per-chunk table limits.
struct CallPlan {
    std::vector<ChunkDesc> desc;
    std::vector<SegInfo> segs;
    std::vector<ChunkReq> creqs;
};

bool planCall(const WavefrontRequest* reqs, uint32_t nreq, uint32_t cap, CallPlan& plan)
{
    uint32_t total = 0, totalPix = 0, nseg = 0, ncreq = 0;
    auto open = [&]() {
        plan.desc.push_back(ChunkDesc{});
        plan.desc.back().segBase = (uint32_t)plan.segs.size();
        plan.desc.back().creqBase = (uint32_t)plan.creqs.size();
        total = totalPix = nseg = ncreq = 0;
    };
    auto close = [&]() { plan.desc.back().hdr = ChunkHdr{total, nseg, ncreq, totalPix}; };
    open();
    for (uint32_t i = 0; i < nreq; i++) {
        const uint32_t count = reqs[i].width * reqs[i].height;
        if (count > cap) return false;
        for (uint32_t s = 0; s < reqs[i].sampleCount; s++) {
            const bool sameReq = ncreq > 0 && plan.creqs.back().req == i;
            if (total + count > cap || nseg == kMaxSegs || (!sameReq && ncreq == kMaxCreqs)) {
                close();
                open();
            }
            if (ncreq == 0 || plan.creqs.back().req != i) {
                plan.creqs.push_back(ChunkReq{i, nseg, 0, totalPix});
                ncreq++;
                totalPix += count;
            }
            plan.creqs.back().nSeg++;
            plan.segs.push_back(SegInfo{i, reqs[i].sampleStart + s, total, count});
            nseg++;
            total += count;
        }
    }
    close();
    if (plan.desc.back().hdr.nseg == 0) plan.desc.pop_back();
    return true;
}

size_t callTableBytes(uint32_t nreq, size_t nchunk, size_t nseg, size_t ncreq)
{
    return kCallHdrBytes + alignUp(sizeof(DevReq) * nreq) + alignUp(sizeof(ChunkDesc) * nchunk) +
           alignUp(sizeof(SegInfo) * nseg) + alignUp(sizeof(ChunkReq) * ncreq);
}

}  // namespace

size_t wavefrontBufferBytes(uint32_t paths) { return alignUp((size_t)paths * kPerPath) + kFixedBytes; }

namespace {

// Device memory of the allocation-free render (aph_render_noalloc): the path state,
// queues and chunk tables of kNoAllocPaths paths in flight, then a call table of
// kNoAllocTableBytes.  Each GPU holds its own copy from the library's first use on it
// until the process ends; every allocation-free call on a GPU works in that GPU's copy.
constexpr size_t kNoAllocPathBytes = ((size_t)kNoAllocPaths * kPerPath + 255) / 256 * 256 + kFixedBytes;
constexpr size_t kNoAllocBytes = kNoAllocPathBytes + kNoAllocTableBytes;
static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

}  // namespace

hipError_t wavefrontScratchState(WavefrontState& st)
{
    // The array's address on the current device, looked up once per device: in host
    // code `scratch` names a host-side stand-in, never the device memory itself.
    static std::mutex mu;
    static std::map<int, void*> perDevice;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    void* ptr = nullptr;
    {
        std::lock_guard<std::mutex> lock(mu);
        auto it = perDevice.find(dev);
        if (it == perDevice.end()) {
            e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
            if (e != hipSuccess) return e;
            perDevice[dev] = ptr;
        } else {
            ptr = it->second;
        }
    }
    st.external = true;
    st.useGraph = false;
    st.buf = ptr;
    st.bytes = kNoAllocPathBytes;
    st.tbl = (uint8_t*)ptr + kNoAllocPathBytes;
    st.tblBytes = kNoAllocTableBytes;
    st.capacity = 0;
    return hipSuccess;
}

size_t wavefrontCallTableBytes(const WavefrontRequest* reqs, uint32_t nreq, uint32_t paths)
{
    size_t nseg = 0;
    uint64_t work = 0;
    for (uint32_t i = 0; i < nreq; i++) {
        nseg += reqs[i].sampleCount;
        work += (uint64_t)reqs[i].width * reqs[i].height * reqs[i].sampleCount;
    }
    // a chunk closes when the next segment does not fit (so two consecutive chunks hold more
    // than the capacity) or at a table limit
    size_t nchunk = 2 + (size_t)(2 * work / std::max<uint32_t>(paths, 1)) + nseg / kMaxSegs + nseg / kMaxCreqs;
    nchunk = std::min(nchunk, nseg + 1);
    return callTableBytes(nreq, nchunk, nseg, nseg);
}

uint32_t wavefrontCapacity(const WavefrontState& st, const WavefrontRequest* reqs, uint32_t nreq, uint32_t chunkPaths)
{
    uint32_t maxPix = 0;
    for (uint32_t i = 0; i < nreq; i++) maxPix = std::max(maxPix, reqs[i].width * reqs[i].height);
    return std::max<uint32_t>(maxPix, chunkPaths ? chunkPaths : st.tuning.chunkPaths);
}

hipError_t wavefrontRender(WavefrontState& st, const WavefrontRequest* reqs, uint32_t nreq, DevCounters* ctr,
                           bool stats, hipStream_t stream, std::vector<PendingEvent>* timing, uint32_t chunkPaths)
{
    if (nreq == 0) return hipSuccess;
    if (nreq >= kCameraTag) return hipErrorInvalidValue;  // request indices travel in 31 bits
    const WavefrontTuning& tn = st.tuning;
    uint32_t cap = wavefrontCapacity(st, reqs, nreq, chunkPaths);
    if (st.external && wavefrontBufferBytes(cap) > st.bytes) {
        // the largest capacity the given memory holds, if it still holds one frame of every request
        const size_t fit = st.bytes > kFixedBytes + 256 ? (st.bytes - kFixedBytes - 256) / kPerPath : 0;
        uint32_t maxPix = 0;
        for (uint32_t i = 0; i < nreq; i++) maxPix = std::max(maxPix, reqs[i].width * reqs[i].height);
        if (fit < maxPix) return hipErrorOutOfMemory;
        cap = (uint32_t)std::min<size_t>(fit, 0xFFFFFFFFu);
    }

    CallPlan plan;
    if (!planCall(reqs, nreq, cap, plan) || plan.desc.empty()) return hipErrorInvalidValue;
    const size_t tblNeed = callTableBytes(nreq, plan.desc.size(), plan.segs.size(), plan.creqs.size());

    // device memory: path state + queues + chunk tables, and the call table
    const size_t need = wavefrontBufferBytes(cap);
    if (st.external) {
        if (need > st.bytes || tblNeed > st.tblBytes) return hipErrorOutOfMemory;
    } else {
        if (st.bytes < need || st.capacity != cap) {
            if (st.buf) {
                hipStreamSynchronize(stream);  // queued chunks still use the old buffers
                if (st.graphExec) hipGraphExecDestroy(st.graphExec);
                st.graphExec = nullptr;
                hipFree(st.buf);
                st.buf = nullptr;
                st.bytes = 0;
            }
            hipError_t e = hipMalloc(&st.buf, need);
            if (e != hipSuccess) {
                st.buf = nullptr;
                st.capacity = 0;
                return e;
            }
            st.bytes = need;
        }
        if (st.tblBytes < tblNeed) {
            if (st.tbl) {
                hipStreamSynchronize(stream);
                if (st.graphExec) hipGraphExecDestroy(st.graphExec);
                st.graphExec = nullptr;
                hipFree(st.tbl);
                st.tbl = nullptr;
                st.tblBytes = 0;
            }
            const size_t grow = std::max(tblNeed, (size_t)1 << 20);
            hipError_t e = hipMalloc(&st.tbl, grow);
            if (e != hipSuccess) {
                st.tbl = nullptr;
                return e;
            }
            st.tblBytes = grow;
        }
    }
    st.capacity = cap;
    char* base = (char*)st.buf;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        void* p = base + off;
        off += alignUp(bytes);
        return p;
    };
    Paths P;
    P.pix = (uint32_t*)take((size_t)cap * 4);
    P.req = (uint32_t*)take((size_t)cap * 4);
    P.rng = (uint2*)take((size_t)cap * 8);
    P.contrib = (float4*)take((size_t)cap * 16);
    P.flags = (uint32_t*)take((size_t)cap * 4);
    P.outDirect = (float4*)take((size_t)cap * 16);
    P.outDiffuse = (float4*)take((size_t)cap * 16);
    P.outGlossy = (float4*)take((size_t)cap * 16);
    P.lastIndirect = (float4*)take((size_t)cap * 16);
    P.pendDirect = (float4*)take((size_t)cap * 16);
    P.pendContrib = (float4*)take((size_t)cap * 16);
    P.nextDir = (float4*)take((size_t)cap * 16);
    P.shColor = (float4*)take((size_t)cap * kSlots * 16);
    P.shMask = (uint8_t*)take((size_t)cap);
    P.vis = (uint8_t*)take((size_t)cap * kSlots);
    P.sampleOut = (float4*)take((size_t)cap * 16);
    Queues Q;
    for (int k = 0; k < 2; k++) {
        Q.rayA[k] = (float4*)take((size_t)cap * 16);
        Q.rayB[k] = (float4*)take((size_t)cap * 16);
    }
    Q.hitA = (float4*)take((size_t)cap * 16);
    Q.hitV = (float*)take((size_t)cap * 4);
    Q.shA = (float4*)take((size_t)cap * kSlots * 16);
    Q.shB = (float4*)take((size_t)cap * kSlots * 16);
    Q.counts = (uint32_t*)take(64);
    Q.cont = (uint8_t*)take((size_t)cap);
    Q.tileCnt = (uint32_t*)take(((size_t)cap + kTile - 1) / kTile * 4);
    Q.tileOff = (uint32_t*)take(((size_t)cap + kTile - 1) / kTile * 4);
    char* ctbl = (char*)take(kTableBytes);
    if (off > need) return hipErrorOutOfMemory;
    ChunkHdr* ghdr = (ChunkHdr*)ctbl;
    SegInfo* gseg = (SegInfo*)(ctbl + kHdrBytes);
    ChunkReq* gcr = (ChunkReq*)(ctbl + kTableBytes / 2);
    uint8_t* tbl = (uint8_t*)st.tbl;
    const DevReq* dreq = (const DevReq*)(tbl + kCallHdrBytes);

    // the call table's host image, uploaded through kernel arguments
    uint32_t maxDepth = 1;
    bool ground = false;
    std::vector<uint32_t> image(tblNeed / 4, 0u);
    uint8_t* img = (uint8_t*)image.data();
    CallHdr ch{};
    ch.nchunk = (uint32_t)plan.desc.size();
    ch.nreq = nreq;
    ch.descOff = (uint32_t)(kCallHdrBytes + alignUp(sizeof(DevReq) * nreq));
    ch.segOff = (uint32_t)(ch.descOff + alignUp(sizeof(ChunkDesc) * plan.desc.size()));
    ch.creqOff = (uint32_t)(ch.segOff + alignUp(sizeof(SegInfo) * plan.segs.size()));
    std::memcpy(img, &ch, sizeof(ch));
    for (uint32_t i = 0; i < nreq; i++) {
        DevReq r;
        std::memset(&r, 0, sizeof(DevReq));
        r.rs = reqs[i].rs;
        r.fp = reqs[i].frame;
        r.W = reqs[i].width;
        r.H = reqs[i].height;
        r.seedOffset = reqs[i].seedOffset;
        r.accum = reqs[i].accum;
        std::memcpy(img + kCallHdrBytes + sizeof(DevReq) * i, &r, sizeof(DevReq));
        maxDepth = std::max<uint32_t>(maxDepth, (uint32_t)reqs[i].frame.trace_depth);
        ground = ground || groundActive(reqs[i].rs.ground);
    }
    std::memcpy(img + ch.descOff, plan.desc.data(), sizeof(ChunkDesc) * plan.desc.size());
    std::memcpy(img + ch.segOff, plan.segs.data(), sizeof(SegInfo) * plan.segs.size());
    std::memcpy(img + ch.creqOff, plan.creqs.data(), sizeof(ChunkReq) * plan.creqs.size());
    hipError_t e = upload(tbl, image, stream);
    if (e != hipSuccess) return e;

    int dev = 0, cus = 256;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    auto grid = [&](uint64_t work, unsigned perCU) {
        return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((work + kWB - 1) / kWB, (uint64_t)cus * perCU));
    };
    Launch L;
    L.gen = grid(cap, 16);
    L.trace = grid(cap, tn.tracePerCU);
    L.shade = grid(cap, tn.shadePerCU);
    L.light = grid(cap, 8);
    L.shadow = grid((uint64_t)cap * 2, tn.tracePerCU);
    L.accum = std::max(1u, std::min<unsigned>((cap + 255) / 256, (unsigned)cus * 8));
    L.refill = tn.refillLanes;
    L.leafWait = tn.leafWaitLanes;

    // One graph per (buffers, capacity, depth, stats, launch sizes): captured on
    // first use, replayed for every chunk.  Timed runs and external states launch
    // eagerly (stage by stage, with HIP events when timed).
    const bool graph = st.useGraph && timing == nullptr && !st.external;
    if (graph) {
        WavefrontState::GraphKey key{st.buf, st.tbl, ctr, cap, maxDepth, stats, ground, L.gen, L.trace, L.shade,
                                     L.shadow, L.accum, L.refill, L.leafWait};
        if (!st.graphExec || !(key == st.graphKey)) {
            if (st.graphExec) hipGraphExecDestroy(st.graphExec);
            st.graphExec = nullptr;
            hipGraph_t g = nullptr;
            if ((e = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal)) != hipSuccess) return e;
            enqueueChunk(stream, L, maxDepth, stats, ground, tbl, dreq, gseg, gcr, ghdr, P, Q, ctr, nullptr);
            if ((e = hipStreamEndCapture(stream, &g)) != hipSuccess) return e;
            e = hipGraphInstantiate(&st.graphExec, g, nullptr, nullptr, 0);
            hipGraphDestroy(g);
            if (e != hipSuccess) {
                st.graphExec = nullptr;
                return e;
            }
            st.graphKey = key;
        }
    }
    // Chunks follow each other on the stream: kChunkNext of chunk k+1 runs after
    // chunk k's last stage, so no host synchronization is needed between them.
    for (size_t k = 0; k < plan.desc.size(); k++) {
        if (graph) {
            if ((e = hipGraphLaunch(st.graphExec, stream)) != hipSuccess) return e;
        } else {
            enqueueChunk(stream, L, maxDepth, stats, ground, tbl, dreq, gseg, gcr, ghdr, P, Q, ctr, timing);
        }
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    return hipSuccess;
}

void wavefrontRelease(WavefrontState& st, hipStream_t stream)
{
    if (st.external) return;
    if (st.buf || st.tbl) hipStreamSynchronize(stream);
    if (st.graphExec) hipGraphExecDestroy(st.graphExec);
    st.graphExec = nullptr;
    if (st.buf) hipFree(st.buf);
    if (st.tbl) hipFree(st.tbl);
    st.buf = nullptr;
    st.tbl = nullptr;
    st.bytes = 0;
    st.tblBytes = 0;
    st.capacity = 0;
}

}  // namespace aph
