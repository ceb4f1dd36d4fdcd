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

// Device buffers kept across calls (grown on demand) and the captured chunk graph.
struct WavefrontState {
    void* buf = nullptr;
    size_t bytes = 0;
    uint32_t capacity = 0;  // paths in flight
    WavefrontTuning tuning;
    bool useGraph = true;
    struct GraphKey {
        void* buf;
        void* ctr;
        uint32_t depth;
        bool stats;
        unsigned gen, trace, shade, shadow, accum, refill, leafWait;
        bool operator==(const GraphKey& o) const
        {
            return buf == o.buf && ctr == o.ctr && depth == o.depth && stats == o.stats && gen == o.gen &&
                   trace == o.trace && shade == o.shade && shadow == o.shadow && accum == o.accum &&
                   refill == o.refill && leafWait == o.leafWait;
        }
    };
    GraphKey graphKey{};
    hipGraphExec_t graphExec = nullptr;
};

hipError_t wavefrontRender(WavefrontState& st, const WavefrontRequest* reqs, uint32_t nreq, DevCounters* ctr,
                           bool stats, hipStream_t stream, std::vector<PendingEvent>* timing);
void wavefrontRelease(WavefrontState& st);

}  // namespace aph
