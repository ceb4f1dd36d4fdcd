// Device-resident scene layout for the gfx950 path tracer.
//
// HBM layout (all arrays 16-byte aligned):
//   BLAS per mesh   BvhNode[n-1]  64 B/node, both children's boxes in the parent
//                   Bvh4Node[n-1] 128 B/node, the traversal copy (four children)
//                   Tri[n]        48 B/triangle, vertices in BVH leaf order,
//                                 original primitive index in v0.w
//   mesh shading    pos float[3V], nrm float[3V] (optional), idx uint[3T]
//   TLAS            BvhNode[m-1]  over instance world boxes, and its Bvh4Node copy
//                   InstTrav[m]   64 B: world->object 3x4 + BLAS pointers
//                   InstShade[m]  object->world 3x4 + shading pointers
//   materials       DevMaterial[k]
//   environment     float4 texels, EnvSample (48 B) per texel
#pragma once

#include "dmath.h"

namespace aph {

// Two-child BVH node: a/b = child 0 box (lo.xyz, ref0) / (hi.xyz, ref1), c/d = child 1 box.
struct alignas(16) BvhNode {
    float4 a, b, c, d;
};

// Child reference: >= 0 internal node index; leaf = kLeafBit | (count-1) << 27 | first.
constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kLeafFirstMask = 0x07FFFFFFu;
constexpr int32_t kExitMarker = 0x7FFFFFFF;  // stack marker: leave the BLAS, back to the TLAS
constexpr uint32_t kMaxLeafPrims = 16;
constexpr int32_t kEmptyChild = 0x7FFFFFFD;  // unused slot of a Bvh4Node (its box is inverted)

// Four-child node, the one traversal reads: node i holds the children of
// two-child node i with every internal child replaced by that child's own two
// children, so indices carry over and the nodes a traversal reaches are those
// at even depths of the two-child tree.  Boxes in SoA form (one float4 per
// plane coordinate, child k in component k); 128 B, one L2 line.
struct alignas(16) Bvh4Node {
    float4 lox, loy, loz, hix, hiy, hiz;
    int4 ref;
    float4 pad;
};
APH_HD int32_t makeLeaf(uint32_t first, uint32_t count)
{
    return (int32_t)(kLeafBit | ((count - 1u) << 27) | first);
}
APH_HD uint32_t leafFirst(int32_t ref) { return (uint32_t)ref & kLeafFirstMask; }
APH_HD uint32_t leafCount(int32_t ref) { return (((uint32_t)ref >> 27) & 0xFu) + 1u; }

struct alignas(16) Tri {
    float4 v0, v1, v2;  // v0.w = original primitive index (bits)
};

struct alignas(16) InstTrav {
    float w2o[12];
    const Bvh4Node* nodes;
    const Tri* tris;
};

struct alignas(16) InstShade {
    float o2w[12];
    const float* pos;
    const float* nrm;
    const float* uv;  // 2 floats per vertex (given, or u = v = i / vertex count)
    const uint32_t* idx;
    uint32_t material;
    uint32_t flags;
    uint32_t layerFirst;  // material layers of the instance: layerMats[layerFirst + k], k < layerCount,
    uint32_t layerCount;  // bottom (k = 0, just above the base) to top
};

struct TraceScene {
    const Bvh4Node* tlas;
    const InstTrav* inst;
    // Shadow transmittance per instance (the DirectX shadow any-hit): xyz = 1 - opacity *
    // (1 - transmission * transmission colour) of the instance's material, w = 1 when the
    // material is opaque (blocks outright).  Read by shadow rays of frames without
    // force_opaque_shadows.
    const float4* transmit;
    float extent;      // 3 x max |coordinate| of all instance world boxes
    float padScale;    // relative pad of instance boxes (times condition number bound)
    uint32_t ninst;
    uint32_t pad_;
};

// Runtime material: the Material struct of Material.slang:20-57; textures in DevMatTex.
struct DevMaterial {
    float base; f3 baseColor;
    float diffuseRoughness, metalness, specular, specularRoughness;
    f3 specularColor; float specularIOR;
    float specularAnisotropy, specularRotation, transmission, subsurface;
    f3 transmissionColor; float sheen;
    f3 subsurfaceColor; float sheenRoughness;
    f3 sheenColor; float coat;
    f3 coatColor; float coatRoughness;
    float coatAnisotropy, coatRotation, coatIOR, coatAffectColor;
    float coatAffectRoughness, emission; int thinWalled; int texMask;  // bit k: slot k's has_*_image flag
    f3 emissionColor; float pad1;
    f3 opacity; float pad2;
};

// Texture slots of a material (in the order of its has_*_image flags).
constexpr int kTexBaseColor = 0, kTexSpecularRoughness = 1, kTexEmissionColor = 2, kTexOpacity = 3,
              kTexNormal = 4, kTexSlots = 5;
constexpr int kAddrWrap = 0, kAddrMirror = 1, kAddrClamp = 2, kAddrMirrorOnce = 3, kAddrBorder = 4;

// An image with its sampler's address modes.
struct alignas(16) DevTexture {
    const float4* texels;  // linear RGBA, row-major
    uint32_t w, h;
    int32_t addrU, addrV;
};

// Texture slots of one material: the bound texture (or -1) and the UV
// transform from the material's texture block (offset.xy, scale.xy, pivot.xy,
// rotation in degrees, pad).
struct DevMatTex {
    int32_t tex[kTexSlots];
    int32_t pad[3];
    float xf[kTexSlots][8];
};

struct alignas(16) AliasEntry {
    uint32_t alias;
    float prob;
    float pdf;
    float pad;
};

// One record per light-image texel for light sampling: the texel's alias-map entry with
// its alias's pdf and both texels' radiance, so a light sample reads one record (three
// 16-byte loads of one line) instead of the entry, possibly its alias's entry, and then a
// texel, each waiting on the last.
struct alignas(16) EnvSample {
    uint32_t alias;
    float prob;
    float pdf, pdfAlias;
    float r, g, b;      // radiance of this texel
    float ra, ga, ba;   // radiance of the alias texel
    float pad[2];
};
static_assert(sizeof(EnvSample) == 48, "three 16-byte loads");

struct DevEnv {
    f3 lightTop, lightBottom, bgTop, bgBottom;
    float lightXf[16], lightXfInv[16], bgXf[16];
    int bgUseScreen, filter, hasLightTex, hasBgTex;
    const float4* lightTex;
    const EnvSample* samples;  // light sampling records, one per light-image texel
    const float4* bgTex;
    uint32_t lw, lh, bw, bh;
    float lightIntegral;
    float pad_;
};

}  // namespace aph
