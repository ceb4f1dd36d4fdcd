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

// Device buffers kept across calls (grown on demand).
struct WavefrontState {
    void* buf = nullptr;
    size_t bytes = 0;
    uint32_t capacity = 0;  // paths in flight
};

hipError_t wavefrontRender(WavefrontState& st, const WavefrontRequest* reqs, uint32_t nreq, DevCounters* ctr,
                           bool stats, hipStream_t stream, std::vector<PendingEvent>* timing);
void wavefrontRelease(WavefrontState& st);

}  // namespace aph
