// Guide images (AOVs) and temporal resolve (temporal.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.h"

namespace aph {

// One plane per guide, one element per pixel (row-major, y down).
struct AovPlanes {
    float* depth;                // view-space depth (negative in front); sky: skyViewZ
    float4* normalRoughness;     // world normal, max(specular roughness, 0.05)
    float4* baseColorMetalness;  // material base colour, metalness
    float4* motion;              // previous - current pixel position; |prev viewZ| - |cur viewZ|
};

struct AovParams {
    aph_frame fp;
    float viewProj[16];      // row-major, clip = M (p, 1)
    float viewProjPrev[16];
    float invViewPrev[16];   // rows right, up, back, eye (the layout of fp.inv_view)
    float skyViewZ;
    uint32_t W, H, sampleIndex;
};

struct TaaParams {
    uint32_t W, H;
    int32_t isReset;
    float alphaMin, alphaMax, speedReference;
};

hipError_t launchAovs(const RenderScene& rs, const AovParams& p, const AovPlanes& out, hipStream_t st);
hipError_t launchTemporalResolve(const float4* accum, const float* depth, const float4* motion, const float4* histIn,
                                 float4* histOut, const TaaParams& p, hipStream_t st);

}  // namespace aph
