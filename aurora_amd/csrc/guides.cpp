// C-ABI of the guide images (AOVs) and the temporal resolve (temporal.hip).
#include <hip/hip_runtime.h>

#include <cstring>

#include "context.h"
#include "temporal.h"

using namespace aph;

namespace {

// halton, PTRenderer.cpp:81-92: radical inverse of index in the given base.
float halton(uint32_t index, uint32_t base)
{
    float result = 0.0f;
    float scale = 1.0f / (float)base;
    while (index > 0) {
        result += scale * (float)(index % base);
        index /= base;
        scale /= (float)base;
    }
    return result;
}

}  // namespace

extern "C" {

void aph_halton_jitter(uint32_t frameIndex, uint32_t phaseCount, float out[2])
{
    if (!out) return;
    uint32_t index = (phaseCount > 0u ? frameIndex % phaseCount : frameIndex) + 1u;
    out[0] = halton(index, 2) - 0.5f;
    out[1] = halton(index, 3) - 0.5f;
}

int aph_aov_create(aph_ctx_t c, uint32_t width, uint32_t height, uint64_t* aov)
{
    int rc = enter(c);
    if (rc) return rc;
    if (!aov || width == 0 || height == 0) return fail(c, APH_E_INVALID, "bad AOV size or null output");
    if ((uint64_t)width * height > (1ull << 28)) return fail(c, APH_E_INVALID, "AOV image too large");
    std::unique_ptr<AovSet> a(new AovSet());
    size_t n = (size_t)width * height;
    a->w = width;
    a->h = height;
    HIPCALL(c, a->depth.alloc(n * sizeof(float)), "alloc AOV depth");
    HIPCALL(c, a->normalRoughness.alloc(n * sizeof(float4)), "alloc AOV normals");
    HIPCALL(c, a->baseColorMetalness.alloc(n * sizeof(float4)), "alloc AOV base colour");
    HIPCALL(c, a->motion.alloc(n * sizeof(float4)), "alloc AOV motion");
    uint64_t h = c->nextHandle++;
    c->aovs[h] = std::move(a);
    *aov = h;
    return APH_OK;
}

int aph_aov_destroy(aph_ctx_t c, uint64_t aov)
{
    int rc = enter(c);
    if (rc) return rc;
    if (!c->aovs.count(aov)) return fail(c, APH_E_INVALID, "unknown AOV set %llu", (unsigned long long)aov);
    HIPCALL(c, hipStreamSynchronize(c->stream), "AOV destroy");
    c->aovs.erase(aov);
    return APH_OK;
}

int aph_aov_readback(aph_ctx_t c, uint64_t aov, float* depth, float* normalRoughness, float* baseColorMetalness,
                     float* motion)
{
    int rc = enter(c);
    if (rc) return rc;
    AovSet* a = lookup(c->aovs, aov);
    if (!a) return fail(c, APH_E_INVALID, "unknown AOV set %llu", (unsigned long long)aov);
    size_t n = (size_t)a->w * a->h;
    HIPCALL(c, hipStreamSynchronize(c->stream), "AOV readback");
    if (depth) HIPCALL(c, hipMemcpy(depth, a->depth.p, n * 4, hipMemcpyDeviceToHost), "AOV readback");
    if (normalRoughness)
        HIPCALL(c, hipMemcpy(normalRoughness, a->normalRoughness.p, n * 16, hipMemcpyDeviceToHost), "AOV readback");
    if (baseColorMetalness)
        HIPCALL(c, hipMemcpy(baseColorMetalness, a->baseColorMetalness.p, n * 16, hipMemcpyDeviceToHost),
                "AOV readback");
    if (motion) HIPCALL(c, hipMemcpy(motion, a->motion.p, n * 16, hipMemcpyDeviceToHost), "AOV readback");
    return APH_OK;
}

int aph_render_aovs(aph_ctx_t c, const aph_aov_request* r)
{
    int rc = enter(c);
    if (rc) return rc;
    if (!r) return fail(c, APH_E_INVALID, "null AOV request");
    if ((rc = checkTextures(c))) return rc;
    AovSet* a = lookup(c->aovs, r->aov);
    Tlas* tl = lookup(c->tlases, r->tlas);
    if (!a) return fail(c, APH_E_INVALID, "unknown AOV set %llu", (unsigned long long)r->aov);
    if (!tl) return fail(c, APH_E_INVALID, "unknown tlas %llu", (unsigned long long)r->tlas);
    AovParams p;
    p.fp = clampFrame(r->frame);
    std::memcpy(p.viewProj, r->view_proj, sizeof(p.viewProj));
    std::memcpy(p.viewProjPrev, r->view_proj_prev, sizeof(p.viewProjPrev));
    std::memcpy(p.invViewPrev, r->inv_view_prev, sizeof(p.invViewPrev));
    p.skyViewZ = r->sky_view_z;
    p.W = a->w;
    p.H = a->h;
    p.sampleIndex = r->sample_index + r->seed_offset;
    AovPlanes out{a->depth.as<float>(), a->normalRoughness.as<float4>(), a->baseColorMetalness.as<float4>(),
                  a->motion.as<float4>()};
    TimedScope ts(c, "aovs");
    HIPCALL(c, launchAovs(renderScene(c, tl), p, out, c->stream), "AOV launch");
    return APH_OK;
}

int aph_temporal_resolve(aph_ctx_t c, const aph_taa_request* r)
{
    int rc = enter(c);
    if (rc) return rc;
    if (!r) return fail(c, APH_E_INVALID, "null temporal resolve request");
    Target* cur = lookup(c->targets, r->current);
    Target* hin = lookup(c->targets, r->history_in);
    Target* hout = lookup(c->targets, r->history_out);
    AovSet* a = lookup(c->aovs, r->aov);
    if (!cur || !hin || !hout || !a) return fail(c, APH_E_INVALID, "unknown target or AOV set");
    if (hin == hout || hout == cur) return fail(c, APH_E_INVALID, "history_out must differ from its inputs");
    for (const Target* t : {hin, hout})
        if (t->w != cur->w || t->h != cur->h) return fail(c, APH_E_INVALID, "history size differs from the frame");
    if (a->w != cur->w || a->h != cur->h) return fail(c, APH_E_INVALID, "AOV size differs from the frame");
    TaaParams p{cur->w, cur->h, r->is_reset, r->blend_alpha_min, r->blend_alpha_max, r->speed_reference};
    TimedScope ts(c, "temporal_resolve");
    HIPCALL(c, launchTemporalResolve(cur->accum.as<float4>(), a->depth.as<float>(), a->motion.as<float4>(),
                                     hin->accum.as<float4>(), hout->accum.as<float4>(), p, c->stream),
            "temporal resolve launch");
    return APH_OK;
}

}  // extern "C"
