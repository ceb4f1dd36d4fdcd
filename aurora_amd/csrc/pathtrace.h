// One pixel-sample of the Aurora path tracer on gfx950: the RayGenShader
// loop (Shaders/MainEntryPoints.slang:70-334) with the closest-hit
// (Shaders/HitShaderEntryPoints.slang:35-486) and miss (:354-426) programs
// inlined and traversal from trace.h.  Shadows are opaque unless the frame turns
// forced-opaque shadows off; the DirectX backend's ground plane (ground.h) and
// material layers apply when the scene has them.
#pragma once

#include "../../include/aph.h"
#include "ground.h"
#include "shade.h"
#include "trace.h"

namespace aph {

struct RenderScene {
    TraceScene sc;
    const InstShade* ishade;
    const DevMaterial* mats;
    const DevMatTex* matTex;      // per material: bound textures and UV transforms
    const DevTexture* textures;
    const uint32_t* layerMats;  // material indices of all instances' layers
    const float* const* layerUV;  // per layer: its geometry's UVs, or null (the instance's); null: none set
    uint32_t nmat;
    uint32_t pad_;
    DevEnv env;
    DevGround ground;
};

// Texture coordinate of a hit (Geometry.slang:199-203): the barycentric mix of
// the vertex UVs, computed on first use (only textured materials read it).  A
// material layer with geometry of its own reads that geometry's UVs at the same
// vertices (computeLayerUV, GlobalBufferAccessors.slang:116-127): source() picks
// the array, null for the instance's own.
struct HitUV {
    const InstShade* ins;
    uint32_t prim;
    float bu, bv;
    bool have;
    const float* src;
    f2 uv;
    __device__ HitUV(const InstShade& i, uint32_t p, float u, float v)
        : ins(&i), prim(p), bu(u), bv(v), have(false), src(nullptr)
    {
    }
    __device__ void source(const float* s)
    {
        if (s != src) {
            src = s;
            have = false;
        }
    }
    __device__ f2 get()
    {
        if (!have) {
            const float* a = src ? src : ins->uv;
            const float b[3] = {1.0f - bu - bv, bu, bv};
            f2 t = f2{0.0f, 0.0f};
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const uint32_t vi = ins->idx[3 * prim + k];
                t = f2{t.x + a[2 * (size_t)vi] * b[k], t.y + a[2 * (size_t)vi + 1] * b[k]};
            }
            uv = t;
            have = true;
        }
        return uv;
    }
};

constexpr uint32_t kNoMaterial = 0xFFFFFFFFu;

// Texture slot of material mi at the hit (an unbound slot reads 0: only the
// emission slot can be flagged and unbound at render time).
__device__ __forceinline__ f3 sampleSlot(const RenderScene& rs, uint32_t mi, int slot, f2 uv)
{
    const DevMatTex& mt = rs.matTex[mi];
    const int t = mt.tex[slot];
    if (t < 0) return splat(0.0f);
    const float4 c = sampleTexture(rs.textures[t], applyUVTransform(uv, mt.xf[slot]));
    return f3{c.x, c.y, c.z};
}

// The material a hit shades with: material layers (HitShaderEntryPoints.slang:84-113,
// DirectX semantics) are tried from the top layer down to the base material; each
// draws its opacity coin (its opacity texture, if flagged, in place of the
// constant), and a layer the coin passes through scales the throughput by
// transparency / P.  Each layer's textures read its own texture coordinates
// (HitShaderEntryPoints.slang:98-111), and the base material the instance's.
// Without layers this is the single coin of the HGI build.
// Returns the hit layer's material index and opacity, or kNoMaterial when every
// layer was passed through.
__device__ __forceinline__ uint32_t selectLayer(const RenderScene& rs, const InstShade& ins, HitUV& huv, Rng& rng,
                                                f3& contrib, float& P, f3& opacity)
{
    for (int li = (int)ins.layerCount; li >= 0; li--) {
        uint32_t mi = li == 0 ? ins.material : rs.layerMats[ins.layerFirst + (uint32_t)li - 1u];
        if (rs.layerUV) huv.source(li == 0 ? nullptr : rs.layerUV[ins.layerFirst + (uint32_t)li - 1u]);
        mi = mi < rs.nmat ? mi : 0u;
        const DevMaterial& m = rs.mats[mi];
        f3 op = m.opacity;
        if (m.texMask & (1 << kTexOpacity)) op = sampleSlot(rs, mi, kTexOpacity, huv.get());
        f3 transparency = splat(1.0f) - op;
        P = luminance(transparency);
        if (random2D(rng).x >= P) {
            opacity = op;
            return mi;
        }
        contrib = contrib * (transparency / P);
    }
    return kNoMaterial;
}

// The textured part of evaluateDefaultMaterial (DefaultMaterial.slang:171-233)
// for the hit material: base colour, specular roughness and emission colour
// into sh / emissionColor, and a normal map replacing the shading normal and
// frame (HitShaderEntryPoints.slang:251-255, buildBasis).
__device__ __forceinline__ void materialTextures(const RenderScene& rs, uint32_t mi, const DevMaterial& m, HitUV& huv,
                                                 Shading& sh, f3& emissionColor)
{
    sh.baseColor = m.baseColor;
    sh.specularRoughness = m.specularRoughness;
    emissionColor = m.emissionColor;
    const int mask = m.texMask;
    if (!mask) return;
    const f2 uv = huv.get();
    if (mask & (1 << kTexBaseColor)) sh.baseColor = sampleSlot(rs, mi, kTexBaseColor, uv);
    if (mask & (1 << kTexSpecularRoughness)) sh.specularRoughness = sampleSlot(rs, mi, kTexSpecularRoughness, uv).x;
    if (mask & (1 << kTexEmissionColor)) emissionColor = sampleSlot(rs, mi, kTexEmissionColor, uv);
    if (mask & (1 << kTexNormal)) {
        sh.normal = normalFromMap(sampleSlot(rs, mi, kTexNormal, uv), sh.normal, sh.tangent);
        buildBasis(sh.normal, sh.tangent, sh.bitangent);
    }
}

struct PathCounters {
    uint32_t ext, shadow;
    TraceStats ts;   // closest-hit queries
    TraceStats tsS;  // any-hit (shadow) queries
};

template <bool STATS>
__device__ __forceinline__ f3 traceShadow(const RenderScene& rs, bool transparent, f3 org, f3 dir,
                                          const StackMem& stk, PathCounters& pc)
{
    pc.shadow++;
    return traceVisibility<STATS>(rs.sc, org, dir, kRayTMin, transparent, stk, pc.tsS);
}

// computeCameraRay, RayTrace.slang:130-182
__device__ __forceinline__ void cameraRay(const aph_frame& fp, uint32_t x, uint32_t y, uint32_t W, uint32_t H, Rng& rng,
                                          f3& org, f3& dir)
{
    float sx = (float)x, sy = (float)y;
    if (fp.temporal_jitter) {
        sx += 0.5f + fp.jitter[0];
        sy += 0.5f + fp.jitter[1];
    } else {
        f2 r = random2D(rng);
        sx += r.x;
        sy += r.y;
    }
    float nx = (sx / (float)W) * 2.0f - 1.0f, ny = (sy / (float)H) * 2.0f - 1.0f;
    ny = -ny;
    f3 right = f3{fp.inv_view[0], fp.inv_view[1], fp.inv_view[2]};
    f3 up = f3{fp.inv_view[4], fp.inv_view[5], fp.inv_view[6]};
    f3 front = -f3{fp.inv_view[8], fp.inv_view[9], fp.inv_view[10]};
    f3 eye = f3{fp.inv_view[12], fp.inv_view[13], fp.inv_view[14]};
    f3 off = right * (fp.view_size[0] * 0.5f * nx) + up * (fp.view_size[1] * 0.5f * ny);
    if (fp.is_ortho) {
        dir = front;
        org = eye + off;
    } else {
        dir = normalize(front + off);
        org = eye;
    }
    if (fp.lens_radius > 0.0f) {
        f3 focal = org + dir * fp.focal_distance;
        f2 r = random2D(rng);
        float rr = fp.lens_radius * sqrtf(r.x), phi = 2.0f * kPi * r.y;
        org = org + right * (rr * cosf(phi)) + up * (rr * sinf(phi));
        dir = normalize(focal - org);
    }
}

// shadeDirectionalLight, ShadeFunctions.slang:51-89
__device__ __forceinline__ f3 shadeDirectional(const aph_distant_light& l, const DevMaterial& m, const Shading& sh, f3 V,
                                               Rng& rng, f3& shadowDir)
{
    f2 r = random2D(rng);
    f3 L = sampleCone(r, f3{l.direction[0], l.direction[1], l.direction[2]}, l.cos_radius);
    f3 f = evalMaterial(m, sh, V, L);
    if (isBlack(f)) {
        shadowDir = splat(0.0f);
        return splat(0.0f);
    }
    shadowDir = L;
    return f3{l.color_intensity[0], l.color_intensity[1], l.color_intensity[2]} * l.color_intensity[3] * f;
}

// The whole path of one pixel-sample; returns RGBA before accumulation. GROUND compiles
// the ground-plane step in: a scene without a ground plane runs the variant without it,
// whose register pressure (and so its spills) the plane's state does not raise.
template <bool STATS, bool GROUND>
__device__ float4 pathSample(const RenderScene& rs, const aph_frame& fp, uint32_t W, uint32_t H, uint32_t px,
                             uint32_t py, uint32_t sampleIndex, const StackMem& stk, PathCounters& pc)
{
    const DevEnv& env = rs.env;
    const int maxDepth = fp.trace_depth;
    // gFrameData.isForceOpaqueShadowsEnabled (HitShaderEntryPoints.slang:42)
    const bool transparentShadows = fp.force_opaque_shadows == 0 && rs.sc.transmit != nullptr;
    Rng rng = rngInit(sampleIndex, W, px, py);
    f3 contrib = splat(1.0f);
    int rayType = kCameraRay, nextRayType = kCameraRay;
    bool isPrimary = true, pathEnded = false;
    float outputAlpha = 0.0f;
    f3 org, dir;
    cameraRay(fp, px, py, W, H, rng, org, dir);
    f3 outDirect = splat(0.0f), outDiffuse = splat(0.0f), outGlossy = splat(0.0f);
    int misLobeID = kNoSecondaryRay;
    f3 direct = splat(0.0f), indirect = splat(0.0f);
    float alpha = 0.0f;
    for (int depth = 0; depth <= maxDepth && !pathEnded; depth++) {
        if (depth == maxDepth) {
            bool bg = isPrimary || rayType == kTransmissionLobe || rayType == kGroundPlaneReflectionRay;
            direct = evalEnvironment(env, dir, bg, px, py, W, H) * contrib;
            pathEnded = true;
            // `indirect` keeps the previous segment's value: the reference does not clear it here.
        } else {
            float minT = depth == 0 ? 0.0f : kRayTMin;
            direct = splat(0.0f);
            indirect = splat(0.0f);
            HitRec hit;
            pc.ext++;
            bool h = traceRay<false, STATS>(rs.sc, org, dir, minT, INFINITY, stk, hit, pc.ts);
            bool skip = false, absorbed = false;
            // evaluateGroundPlane (GroundPlane.slang:232-290), first in the hit and miss programs
            f3 gpos;
            float gdist;
            if (GROUND && isPrimary && rs.ground.enabled && groundPlaneHit(rs.ground, org, dir, gpos, gdist) &&
                gdist < (h ? (float)hit.t : INFINITY)) {
                f3 gcol, gdir, sdir;
                bool gshadow;
                const int gt = groundPlaneSetup(rs.ground, env, dir, rng, gcol, gdir, gshadow, sdir);
                if (gshadow) gcol = gcol * groundShadowTint(rs.ground, traceShadow<STATS>(rs, false, gpos, sdir, stk, pc));
                contrib = contrib * gcol;
                if (gt == kGroundPlaneReflectionRay) {
                    // the segment ends at the plane; the path goes on as its reflection
                    org = gpos;
                    dir = gdir;
                    nextRayType = kGroundPlaneReflectionRay;
                    isPrimary = false;
                    absorbed = true;
                }
            }
            if (absorbed) {
            } else if (!h) {
                bool bg = isPrimary || rayType == kTransmissionLobe || rayType == kGroundPlaneReflectionRay;
                if (bg) direct = evalEnvironment(env, dir, true, px, py, W, H) * contrib;
                pathEnded = true;
            } else {
                // ---- closest hit ----
                f3 nextContrib = splat(1.0f);
                f3 V = -dir;
                const InstShade& ins = rs.ishade[hit.inst];
                Shading sh;
                computeShading(ins, hit.prim, (float)hit.u, (float)hit.v, sh);
                HitUV huv(ins, hit.prim, (float)hit.u, (float)hit.v);
                float P;
                f3 opacity;
                const uint32_t mi = selectLayer(rs, ins, huv, rng, contrib, P, opacity);
                if (mi != kNoMaterial) {
                    const DevMaterial& m = rs.mats[mi];
                    f3 emissionColor;
                    materialTextures(rs, mi, m, huv, sh, emissionColor);
                    direct = direct + emissionColor * m.emission * lerp(splat(1.0f), m.coatColor, m.coat);
                    f3 dirShadow = splat(0.0f), dirColor = splat(0.0f);
                    bool hasDir = false;
                    if (fp.light_count > 1) {
                        f3 o2 = dot(V, sh.normal) < 0.0f ? sh.geomPosition : sh.position;
                        for (int li = 0; li < fp.light_count; li++) {
                            if (fp.lights[li].color_intensity[3] <= 0.0f) continue;
                            f3 sd;
                            f3 lr = shadeDirectional(fp.lights[li], m, sh, V, rng, sd);
                            if (isBlack(lr)) continue;
                            direct = direct + traceShadow<STATS>(rs, transparentShadows, o2, sd, stk, pc) * lr;
                        }
                    } else if (fp.light_count > 0) {
                        // one light: the selection still draws a random number (HitShaderEntryPoints.slang:322-341)
                        const aph_distant_light& l0 = fp.lights[0];
                        float w0 = luminance(f3{l0.color_intensity[0], l0.color_intensity[1], l0.color_intensity[2]}) *
                                   l0.color_intensity[3];
                        float selPdf = 1.0f;
                        random2D(rng);
                        if (w0 > 0.0f) selPdf = w0 / w0;
                        if (l0.color_intensity[3] > 0.0f) {
                            dirColor = shadeDirectional(l0, m, sh, V, rng, dirShadow) / fmaxf(selPdf, 1e-4f);
                            hasDir = true;
                        }
                    }
                    // ---- environment MIS, ShadeFunctions.slang:135-259 ----
                    bool emitsMat = false, emitsLight = false;
                    int misLobe = kNoSecondaryRay;
                    f3 matRes = splat(0.0f), matDir = splat(0.0f), lightRes = splat(0.0f), lightDir = splat(0.0f);
                    {
                        f2 r = random2D(rng);
                        f3 L;
                        float matPdf;
                        int lobe;
                        f3 f = sampleMaterial(m, sh, V, r, L, matPdf, lobe);
                        misLobe = lobe;
                        if (lobe != kTransmissionLobe) {
                            f3 Le = evalEnvironment(env, L, false, px, py, W, H);
                            if (!isBlack(Le) && !isBlack(f)) {
                                float lightPdf = env.hasLightTex ? luminance(Le) / env.lightIntegral : 0.25f * kPiInv;
                                matRes = Le * misWeight(matPdf, lightPdf) * f / matPdf;
                                matDir = L;
                                emitsMat = true;
                            }
                            float lightPdf;
                            Le = sampleEnvironment(env, r, L, lightPdf);
                            int evLobe;
                            f = evalMaterialPdf(m, sh, V, L, r, matPdf, evLobe);
                            if (!isBlack(Le) && !isBlack(f)) {
                                lightRes = Le * misWeight(lightPdf, matPdf) * f / lightPdf;
                                lightDir = L;
                                emitsLight = true;
                            }
                        }
                    }
                    contrib = contrib * (opacity / (1.0f - P));
                    outputAlpha = (outputAlpha == 0.0f && rayType == kTransmissionLobe) ? 0.0f : 1.0f;
                    // ---- setupIndirectLightRay, ShadeFunctions.slang:263-310 ----
                    {
                        f2 r = random2D(rng);
                        f3 L;
                        float pdf;
                        int lobe;
                        f3 f = sampleMaterial(m, sh, V, r, L, pdf, lobe);
                        if (isBlack(f)) {
                            nextRayType = kNoSecondaryRay;
                        } else {
                            bool trans = lobe == kTransmissionLobe;
                            org = (dot(V, sh.normal) < 0.0f || trans) ? sh.geomPosition : sh.position;
                            dir = L;
                            nextContrib = f / pdf;
                            nextRayType = lobe;
                        }
                    }
                    f3 so = dot(V, sh.normal) < 0.0f ? sh.geomPosition : sh.position;
                    if (hasDir) direct = direct + traceShadow<STATS>(rs, transparentShadows, so, dirShadow, stk, pc) * dirColor;
                    if (emitsMat || emitsLight) {
                        f3 mis = splat(0.0f);
                        if (emitsMat) mis = mis + traceShadow<STATS>(rs, transparentShadows, so, matDir, stk, pc) * matRes;
                        if (emitsLight) mis = mis + traceShadow<STATS>(rs, transparentShadows, so, lightDir, stk, pc) * lightRes;
                        indirect = mis;
                        misLobeID = misLobe;
                    }
                    if (nextRayType == kNoSecondaryRay) pathEnded = true;
                    isPrimary = false;
                } else {
                    // every layer passed through: continue straight on, as if nothing had been hit
                    dir = -V;
                    org = sh.geomPosition;
                    skip = true;
                }
                if (!skip) {
                    direct = direct * contrib;
                    indirect = indirect * contrib;
                    contrib = contrib * nextContrib;
                    if (fp.rr_enabled && depth + 1 >= fp.rr_start_depth) {
                        float surv = clampf(maxc(contrib), 0.05f, 1.0f);
                        if (random2D(rng).x > surv) pathEnded = true;
                        else contrib = contrib / surv;
                    }
                }
            }
            alpha = outputAlpha;
            if (skip) continue;
        }
        // Route into the direct / diffuse / glossy outputs like MainEntryPoints.slang:160-243.
        if (depth == 0) {
            isPrimary = false;
            outDirect = direct;
            if (misLobeID == kDiffuseLobe) outDiffuse = indirect;
            else outGlossy = indirect;
        } else {
            f3 radiance = direct + indirect;
            if (rayType == kDiffuseLobe) outDiffuse = outDiffuse + radiance;
            else outGlossy = outGlossy + radiance;
        }
        rayType = nextRayType;
    }
    f3 fin = adjustRadiance(fp.max_luminance, fp.display_errors != 0, outDirect + outDiffuse + outGlossy);
    return make_float4(fin.x, fin.y, fin.z, alpha);
}

}  // namespace aph
