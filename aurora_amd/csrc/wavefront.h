// Wavefront path-tracing pipeline (wavefront.hip): raygen, extend, shade,
// shadow and accumulate stages over SoA path queues shared by many requests.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "kernels.h"
#include "timing.h"

namespace aph {

struct WavefrontRequest {
    RenderScene rs;
    aph_frame frame;
    uint32_t width, height;
    uint32_t sampleStart, sampleCount, seedOffset;
    float4* accum;
};

// Launch parameters of the pipeline (aph_wavefront_autotune picks them).
struct WavefrontTuning {
    uint32_t chunkPaths = 4u << 20;  // paths in flight per chunk
    uint32_t tracePerCU = 16;        // blocks per CU of the persistent traversal kernels
    uint32_t shadePerCU = 8;         // blocks per CU of the shade kernel
    uint32_t refillLanes = 32;       // idle lanes of a traversal wave that trigger a refill from the queue
    uint32_t leafWaitLanes = 16;     // lanes still looking for a leaf when a traversal wave processes leaves
};

// Device buffers kept across calls and the captured chunk graph.
//   buf: path state and queues of `capacity` paths, then the tables of the chunk
//        in flight (fixed addresses, so one captured graph serves every chunk);
//   tbl: the call table - request records and every chunk's tables - uploaded
//        once per call.
// An external state (the allocation-free entry point) works in memory it was
// given and never allocates, frees or captures a graph.
struct WavefrontState {
    void* buf = nullptr;
    size_t bytes = 0;
    uint32_t capacity = 0;  // paths in flight
    void* tbl = nullptr;
    size_t tblBytes = 0;
    bool external = false;
    WavefrontTuning tuning;
    bool useGraph = true;
    struct GraphKey {
        void* buf;
        void* tbl;
        void* ctr;
        uint32_t capacity;
        uint32_t depth;
        bool stats;
        bool ground;  // a request has a ground plane: the chunk runs its stage
        unsigned gen, trace, shade, shadow, accum, refill, leafWait;
        bool operator==(const GraphKey& o) const
        {
            return buf == o.buf && tbl == o.tbl && ctr == o.ctr && capacity == o.capacity && depth == o.depth &&
                   stats == o.stats && ground == o.ground && gen == o.gen && trace == o.trace && shade == o.shade &&
                   shadow == o.shadow && accum == o.accum && refill == o.refill && leafWait == o.leafWait;
        }
    };
    GraphKey graphKey{};
    hipGraphExec_t graphExec = nullptr;
};

// Device bytes of the path state, queues and chunk tables for `paths` paths in flight.
size_t wavefrontBufferBytes(uint32_t paths);
// Bytes of the call table for nreq requests cut into chunks of at most `paths`
// paths (an upper bound computed from the requests alone).
size_t wavefrontCallTableBytes(const WavefrontRequest* reqs, uint32_t nreq, uint32_t paths);
// Paths in flight of a render of these requests: the tuned chunk (or
// chunkPaths when nonzero), and at least one full frame of the largest request.
uint32_t wavefrontCapacity(const WavefrontState& st, const WavefrontRequest* reqs, uint32_t nreq,
                           uint32_t chunkPaths = 0);

// Renders the requests.  Asynchronous on `stream`: no host memory is referenced
// after the call returns.  chunkPaths (nonzero) overrides the tuned chunk size.
// An external state fails with hipErrorOutOfMemory when the render does not
// fit its memory.
hipError_t wavefrontRender(WavefrontState& st, const WavefrontRequest* reqs, uint32_t nreq, DevCounters* ctr,
                           bool stats, hipStream_t stream, std::vector<PendingEvent>* timing,
                           uint32_t chunkPaths = 0);
// The allocation-free render's limits: paths in flight and call-table bytes.
constexpr uint32_t kNoAllocPaths = 1u << 17;
constexpr size_t kNoAllocTableBytes = 1u << 20;
// Points an external state at this device's allocation-free render memory (looked up
// once per device; the device must be current).
hipError_t wavefrontScratchState(WavefrontState& st);
// Frees the buffers of a state that owns them (waits for `stream` first).
void wavefrontRelease(WavefrontState& st, hipStream_t stream = nullptr);

}  // namespace aph
