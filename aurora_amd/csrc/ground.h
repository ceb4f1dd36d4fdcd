// The ground plane of the DirectX backend (Shaders/GroundPlane.slang:24-290): a
// matte plane evaluated on primary rays only, in front of whatever the ray hit.
// A stochastic coin of probability reflection_opacity turns the ray into a GGX-Smith
// reflection off the plane (the path continues from the plane); otherwise the ray
// goes on to its hit or miss with its contribution tinted by a matte shadow: one
// environment sample, one opaque shadow ray from the plane, and the shadow colour
// blended in by the luminance of the blocked light times shadow_opacity.
#pragma once

#include "shade.h"

namespace aph {

// GroundPlane (GlobalPipelineState.slang:57-70); tangent and bitangent come from the
// host's basis of the normal (PTGroundPlane.cpp buildBasis).
struct DevGround {
    int enabled;
    f3 position, normal, tangent, bitangent;
    float shadowOpacity;
    f3 shadowColor;
    float reflectionOpacity;
    f3 reflectionColor;
    float reflectionRoughness;
};

// A plane with neither a reflection nor a shadow changes nothing: its setup draws no
// random number and scales the path by exactly 1.  Such a plane runs the kernels of a
// scene without one (the render kernels are compiled with and without the plane's
// step, and the two compilations need not round alike).
inline bool groundActive(const DevGround& g)
{
    return g.enabled != 0 && (g.reflectionOpacity > 0.0f || g.shadowOpacity > 0.0f);
}

// intersectGroundPlane (GroundPlane.slang:104-128): the plane faces its normal; rays
// travelling along it, or with the plane behind them, miss.
APH_HD bool groundPlaneHit(const DevGround& g, f3 org, f3 dir, f3& pos, float& dist)
{
    const float planeDot = dot(dir, g.normal);
    if (planeDot >= 0.0f) return false;
    dist = dot(g.position - org, g.normal) / planeDot;
    if (dist < 0.0f) return false;
    pos = org + dir * dist;
    return true;
}

// sampleGGXSmithBRDF (BSDFCommon.slang:331-367) with anisotropy 0, in tangent space: the
// BRDF value for the sampled L (black when L leaves V's hemisphere) and
// pdf = D G1 (V.H) / |V.z| / (4 V.H), from the BSDFCommon helpers of shade.h.
APH_HD f3 sampleGgxSmithBrdf(f3 V, f3 f0, float roughness, f2 r, f3& L, float& pdf)
{
    const float a = fmaxf(0.0001f, sqr(roughness));  // roughnessToAlpha, :201-210
    const bool flip = V.z <= 0.0f;
    f3 H = sampleVndf(flip ? -V : V, a, a, r);
    H = flip ? -H : H;
    L = reflect(-V, H);
    pdf = ggxVndf(V, H, a, a) / (4.0f * dot(V, H));
    if (!(V.z * L.z > 0.0f)) return splat(0.0f);
    // evaluateGGXSmithBRDF (:303-323) at the half vector of V and L
    const f3 Hr = normalize(V + L);
    const float D = ggxNdf(Hr, a, a);
    float G2 = 0.0f;
    if (!(dot(V, Hr) <= 0.0f || dot(L, Hr) < 0.0f)) G2 = 1.0f / (1.0f + smithLambda(V, a, a) + smithLambda(L, a, a));
    const f3 F = lerp(f0, splat(1.0f), powf(1.0f - dot(V, Hr), 5.0f));
    return F * (D * G2) / (4.0f * fabsf(V.z) * fabsf(L.z));
}

// setupGroundPlaneRay (GroundPlane.slang:133-193) for a primary ray hitting the plane at
// pos.  Returns kGroundPlaneReflectionRay with the reflection direction in nDir (the ray
// leaves from pos), or kCameraRay (plane only tints the ray; with `shadow` set the caller
// traces an opaque shadow ray along shadowDir from pos and multiplies `color` by
// groundShadowTint of its visibility), or kNoSecondaryRay (a black reflection sample: the
// ray goes on, tinted, without a shadow).  `color` scales the path's contribution.
__device__ __forceinline__ int groundPlaneSetup(const DevGround& g, const DevEnv& env, f3 dir, Rng& rng, f3& color,
                                                f3& nDir, bool& shadow, f3& shadowDir)
{
    color = splat(1.0f);
    shadow = false;
    if (g.reflectionOpacity > 0.0f) {
        const f3 transparency = splat(1.0f - g.reflectionOpacity);
        const float Pr = luminance(transparency);
        if (random2D(rng).x >= Pr) {
            // setupMatteReflectionIndirectLightRay (:24-65)
            const f3 V = f3{dot(-dir, g.tangent), dot(-dir, g.bitangent), dot(-dir, g.normal)};
            f3 L;
            float pdf;
            const f2 r = random2D(rng);
            f3 bc = sampleGgxSmithBrdf(V, g.reflectionColor, g.reflectionRoughness, r, L, pdf);
            bc = bc * fabsf(L.z);
            int type = kNoSecondaryRay;
            if (!isBlack(bc)) {
                nDir = g.tangent * L.x + g.bitangent * L.y + g.normal * L.z;
                color = bc / pdf;
                type = kGroundPlaneReflectionRay;
            }
            color = color * (g.reflectionOpacity / (1.0f - Pr));
            return type;
        }
        color = transparency / Pr;
    }
    if (g.shadowOpacity > 0.0f) {
        // shadeMatteShadow (:67-101): a light below the plane casts no shadow on it
        float pdf;
        f3 L;
        sampleEnvironment(env, random2D(rng), L, pdf);
        if (!(dot(L, g.normal) < 0.0f)) {
            shadow = true;
            shadowDir = L;
        }
    }
    return kCameraRay;
}

// The matte shadow's tint for a shadow ray of visibility vis (:92-100, :181-188).
APH_HD f3 groundShadowTint(const DevGround& g, f3 vis)
{
    const float a = luminance(splat(1.0f) - vis) * g.shadowOpacity;
    return lerp(splat(1.0f), g.shadowColor, a);
}

}  // namespace aph
