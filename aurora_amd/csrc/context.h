// Internal state of a backend context, shared by the C-ABI translation units
// (api.cpp: scene objects and rendering; server.cpp: the serving loop).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <map>
#include <set>
#include <memory>
#include <string>
#include <vector>

#include "../../include/aph.h"
#include "bvh_build.h"
#include "kernels.h"
#include "timing.h"
#include "wavefront.h"

struct aph_denoiser;

namespace aph {

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes)
    {
        o.p = nullptr;
        o.bytes = 0;
    }
    DevBuf& operator=(DevBuf&& o) noexcept
    {
        if (this != &o) {
            reset();
            p = o.p;
            bytes = o.bytes;
            o.p = nullptr;
            o.bytes = 0;
        }
        return *this;
    }
    ~DevBuf() { reset(); }
    void reset()
    {
        if (p) hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    hipError_t alloc(size_t n)
    {
        reset();
        if (n == 0) n = 16;
        hipError_t e = hipMalloc(&p, n);
        if (e == hipSuccess) bytes = n;
        else p = nullptr;
        return e;
    }
    // Grows (never shrinks) a scratch buffer.
    hipError_t reserve(size_t n)
    {
        if (bytes >= n && p) return hipSuccess;
        return alloc(n);
    }
    template <class T> T* as() const { return (T*)p; }
};

struct Mesh {
    uint32_t nv = 0, nt = 0;
    bool hasNrm = false;
    DevBuf pos, nrm, uv, idx, nodes, nodes4, tris;
    bool built = false;
    BuildInfo info{};
};

// A lat-long environment resident on the device: light image + alias map,
// background image, and the DevEnv record that points at them.
struct Environment {
    DevBuf light, samples, bg;  // light image, light sampling records (EnvSample), background
    DevEnv env{};
    size_t bytes() const { return light.bytes + samples.bytes + bg.bytes; }
};

struct Tlas {
    uint32_t n = 0, nvalid = 0;
    const Environment* environment = nullptr;  // aph_tlas_set_environment; null: the context's
    DevBuf nodes, nodes4, trav, shade;
    TraceScene sc{};
    DevBuf layerMats;  // material indices of the instances' layers (aph_tlas_set_layers)
    DevBuf layerUV;    // per layer: device pointer to its geometry's UVs, or null (aph_tlas_set_layer_geometry)
    std::vector<InstShade> shadeHost;
    std::vector<uint32_t> remap;  // indices of the instances with geometry (those the tree holds), ascending
    uint32_t depth = 0;
    uint32_t nodeCount = 0;       // two-child nodes of the tree
    DevGround ground{};           // aph_tlas_set_ground_plane (enabled = 0: none)
    DevBuf transmit;              // shadow transmittance per instance (TraceScene::transmit)
    uint64_t transmitVersion = 0; // the material table version it was computed from (0: never)
    size_t bytes = 0;  // device memory of the TLAS and the meshes/BLASes it references
};

struct Target {
    uint32_t w = 0, h = 0;
    DevBuf accum;
};

struct Image {
    uint32_t w = 0, h = 0;
    DevBuf texels;  // linear RGBA float
};

// sRGB transfer function undone (IEC 61966-2-1), the decode of an sRGB texture format.
inline float srgbToLinear(float c)
{
    return c <= 0.04045f ? c / 12.92f : std::pow((c + 0.055f) / 1.055f, 2.4f);
}

struct AovSet {
    uint32_t w = 0, h = 0;
    DevBuf depth, normalRoughness, baseColorMetalness, motion;
};

struct Timing {
    double ms = 0.0;
    uint64_t launches = 0;
};

struct DenoiserDeleter {
    void operator()(aph_denoiser* d) const;  // denoiser.cpp
};

}  // namespace aph

// The opaque aph_ctx_t of the C-ABI.
struct aph_ctx {
    using DevBuf = aph::DevBuf;
    using Mesh = aph::Mesh;
    using Tlas = aph::Tlas;
    using Target = aph::Target;
    using Timing = aph::Timing;
    using AovSet = aph::AovSet;
    using Image = aph::Image;
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    uint64_t nextHandle = 1;
    std::map<uint64_t, std::unique_ptr<Mesh>> meshes;
    std::map<uint64_t, std::unique_ptr<Tlas>> tlases;
    std::map<uint64_t, std::unique_ptr<Target>> targets;
    std::map<uint64_t, std::unique_ptr<AovSet>> aovs;
    std::map<uint64_t, std::unique_ptr<aph_denoiser, aph::DenoiserDeleter>> denoisers;
    DevBuf mats;
    uint32_t nmat = 0;
    uint64_t matVersion = 1;                       // bumped by every change of the material table
    std::vector<aph_material> matsHost;            // the uploaded table (texture flags and transforms)
    std::vector<aph_texture_binding> texBindings;  // APH_TEX_SLOTS per material, or empty
    std::vector<uint32_t> texturesMissing;         // material * APH_TEX_SLOTS + slot: flagged, unbound
    DevBuf matTex, textures;                       // DevMatTex per material, DevTexture table
    std::map<uint64_t, std::unique_ptr<Image>> images;
    std::unique_ptr<aph::Environment> env;  // aph_environment_upload: scenes without their own
    std::map<uint64_t, std::unique_ptr<aph::Environment>> environments;
    DevBuf counters;
    int statsLevel = 0;
    DevBuf scratchA, scratchB;
    aph::WavefrontState wf;
    aph::WavefrontState wfStatic;   // aph_render_noalloc: the device's scratch array, set up by aph_create
    int wfStaticStatus = -1;        // hipError_t of that set-up (hipSuccess = 0)
    bool timing = false;
    std::map<std::string, Timing> timings;
    std::vector<aph::PendingEvent> pending;
};

namespace aph {

inline int fail(aph_ctx* c, int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

inline int hipFail(aph_ctx* c, hipError_t e, const char* what)
{
    return fail(c, e == hipErrorOutOfMemory ? APH_E_MEMORY : APH_E_DEVICE, "%s: %s", what, hipGetErrorString(e));
}

#define HIPCALL(ctx, x, what)                              \
    do {                                                   \
        hipError_t e_ = (x);                               \
        if (e_ != hipSuccess) return hipFail(ctx, e_, what); \
    } while (0)

template <class T> T* lookup(std::map<uint64_t, std::unique_ptr<T>>& m, uint64_t h)
{
    auto it = m.find(h);
    return it == m.end() ? nullptr : it->second.get();
}

// Frame parameters as RendererBase::updateFrameDataGPUStruct bounds them:
// traceDepth in [1, 10] (RendererBase.cpp:265), RR start depth >= 0 (:274-275).
inline aph_frame clampFrame(const aph_frame& in)
{
    aph_frame fp = in;
    fp.trace_depth = fp.trace_depth < 1 ? 1 : (fp.trace_depth > 10 ? 10 : fp.trace_depth);
    if (fp.rr_start_depth < 0) fp.rr_start_depth = 0;
    return fp;
}

// Everything a render of one TLAS reads: geometry, instance records,
// the context's material table and environment.
inline RenderScene renderScene(const aph_ctx* c, const Tlas* tl)
{
    RenderScene rs;
    rs.sc = tl->sc;
    rs.ishade = tl->shade.as<InstShade>();
    rs.mats = c->mats.as<DevMaterial>();
    rs.matTex = c->matTex.as<DevMatTex>();
    rs.textures = c->textures.as<DevTexture>();
    rs.nmat = c->nmat;
    rs.layerMats = tl->layerMats.as<uint32_t>();
    rs.layerUV = tl->layerUV.as<const float*>();
    rs.pad_ = 0;
    rs.env = tl->environment ? tl->environment->env : c->env->env;
    rs.sc.transmit = tl->transmit.as<float4>();
    rs.ground = tl->ground;
    return rs;
}

// The per-instance shadow transmittance of a TLAS for the current material table
// (the shadow any-hit shader's opacity, HitShaderEntryPoints.slang:489-553; opaque as
// MaterialBase::isOpaque decides: opacity 1, no opacity image, no transmission), recomputed
// when the table has changed since.  Material textures are not sampled for shadows: an
// instance with an opacity image transmits by its constants.
inline int ensureTransmit(aph_ctx* c, Tlas* tl)
{
    if (tl->transmitVersion == c->matVersion && tl->transmit.p) return APH_OK;
    const size_t n = tl->shadeHost.size();
    std::vector<float4> tm(n ? n : 1, make_float4(1.0f, 1.0f, 1.0f, 0.0f));
    for (size_t i = 0; i < n; i++) {
        uint32_t mi = tl->shadeHost[i].material;
        if (mi >= c->nmat) mi = 0;  // as the shading of an out-of-range index
        const aph_material& m = c->matsHost[mi];
        const bool opacityImage = m.texture_block[25] != 0;  // has_opacity_image (the opacity slot's flag word)
        const bool opaque = m.opacity[0] == 1.0f && m.opacity[1] == 1.0f && m.opacity[2] == 1.0f &&
                            m.transmission == 0.0f && !opacityImage;
        float t[3];
        for (int k = 0; k < 3; k++) t[k] = 1.0f - m.opacity[k] * (1.0f - m.transmission * m.transmission_color[k]);
        tm[i] = make_float4(t[0], t[1], t[2], opaque ? 1.0f : 0.0f);
    }
    HIPCALL(c, hipStreamSynchronize(c->stream), "shadow transmittance");  // renders in flight read the old one
    HIPCALL(c, tl->transmit.alloc(sizeof(float4) * tm.size()), "alloc shadow transmittance");
    HIPCALL(c, hipMemcpy(tl->transmit.p, tm.data(), sizeof(float4) * tm.size(), hipMemcpyHostToDevice),
            "copy shadow transmittance");
    tl->transmitVersion = c->matVersion;
    return APH_OK;
}

// A render reads every flagged texture slot; an unbound one (other than
// emission, which reads 0) cannot be sampled.
inline int checkTextures(aph_ctx* c)
{
    if (c->texturesMissing.empty()) return APH_OK;
    const uint32_t k = c->texturesMissing[0];
    return fail(c, APH_E_STATE, "material %u: texture slot %u is flagged in the material but has no bound image",
                k / APH_TEX_SLOTS, k % APH_TEX_SLOTS);
}

inline int enter(aph_ctx* c)
{
    if (!c) return APH_E_INVALID;
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return hipFail(c, e, "hipSetDevice");
    c->err.clear();
    return APH_OK;
}

inline void resolveTimings(aph_ctx* c)
{
    for (auto& p : c->pending) {
        hipEventSynchronize(p.b);
        float ms = 0.0f;
        hipEventElapsedTime(&ms, p.a, p.b);
        Timing& t = c->timings[p.name];
        t.ms += ms;
        t.launches++;
        hipEventDestroy(p.a);
        hipEventDestroy(p.b);
    }
    c->pending.clear();
}

struct TimedScope {
    aph_ctx* c;
    const char* name;
    hipEvent_t a = nullptr, b = nullptr;
    TimedScope(aph_ctx* ctx, const char* n) : c(ctx), name(n)
    {
        if (c->timing) {
            hipEventCreate(&a);
            hipEventCreate(&b);
            hipEventRecord(a, c->stream);
        }
    }
    ~TimedScope()
    {
        if (c->timing) {
            hipEventRecord(b, c->stream);
            c->pending.push_back(PendingEvent{name, a, b});
        }
    }
};

}  // namespace aph
