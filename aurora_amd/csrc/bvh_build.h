// BVH builder interfaces (bvh.hip: LBVH; sah.hip: binned SAH).
#pragma once

#include <hip/hip_runtime.h>

#include "scene.h"

namespace aph {

struct BuildInfo {
    float rootLo[3], rootHi[3];
    uint32_t depth;  // longest root-to-leaf path in nodes (an upper bound for SAH)
    uint32_t nodes;  // node records written (root at 0)
};

// Builds max(n-1, 1) nodes into `nodes` from primitive boxes.  sortedOut[i] is
// the primitive stored at leaf position i.  leafOriginal: leaf references carry
// the original primitive index (TLAS) instead of the leaf position (BLAS),
// mapped through leafRemap when it is non-null.
hipError_t lbvhBuild(const float4* primLo, const float4* primHi, uint32_t n, bool leafOriginal,
                     const uint32_t* leafRemap, hipStream_t st, BvhNode* nodes, uint32_t* sortedOut,
                     BuildInfo* info);
// Binned-SAH build (sah.hip) with the same output contract as lbvhBuild; leaves
// hold up to maxLeaf primitives (1 when leafOriginal).
hipError_t sahBuild(const float4* primLo, const float4* primHi, uint32_t n, uint32_t maxLeaf, bool leafOriginal,
                    const uint32_t* leafRemap, hipStream_t st, BvhNode* nodes, uint32_t* sortedOut,
                    BuildInfo* info);
// Writes the four-child traversal copy of an n-node two-child tree (n >= 1).
hipError_t collapse4(const BvhNode* nodes, uint32_t n, hipStream_t st, Bvh4Node* out);
hipError_t triBoxes(const float* pos, const uint32_t* idx, uint32_t n, hipStream_t st, float4* lo, float4* hi);
hipError_t reorderTris(const float* pos, const uint32_t* idx, const uint32_t* sorted, uint32_t n, hipStream_t st,
                       Tri* tris);

}  // namespace aph
