// Host-side launchers of the path-tracing kernels (kernels.hip, wavefront.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/aph.h"
#include "pathtrace.h"

namespace aph {

struct DevCounters {
    unsigned long long ext, shadow, paths, nodes, tris, insts, overflow;
    unsigned long long snodes, stris, sinsts;  // the any-hit (shadow) share of nodes/tris/insts
    unsigned long long hits;                   // closest-hit queries that found a triangle (wavefront, stats mode)
};

hipError_t launchRenderMega(const RenderScene& rs, const aph_frame& fp, uint32_t W, uint32_t H, uint32_t start,
                            uint32_t count, uint32_t seedOffset, float4* accum, DevCounters* ctr, bool stats,
                            hipStream_t st);
hipError_t launchTrace(const TraceScene& sc, const aph_ray* rays, uint32_t n, aph_hit* hits, uint8_t* occ,
                       DevCounters* ctr, bool any, bool stats, hipStream_t st);
hipError_t launchPost(const float4* accum, uint32_t n, const float brightness[3], int toneMap, int gamma, float4* out,
                      hipStream_t st);

}  // namespace aph
