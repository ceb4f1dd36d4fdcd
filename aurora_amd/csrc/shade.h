// Device shading for the gfx950 path tracer: RNG, sampling, environment,
// Standard Surface BSDF and shading-data interpolation.  Follows the reference
// shaders (paths relative to Libraries/Aurora/Source/Shaders/).
#pragma once

#include "scene.h"

namespace aph {

// Lobe IDs / ray types, BSDFCommon.slang:28-39.
constexpr int kCoatLobe = 2, kMetalLobe = 3, kGlossyLobe = 4, kTransmissionLobe = 5, kSheenLobe = 6,
              kDiffuseLobe = 7, kNoSecondaryRay = 8, kCameraRay = 9, kGroundPlaneReflectionRay = 10;

// ---------------- RNG: PCG2D, Random.slang:96-146 ----------------
struct Rng {
    uint32_t x, y;
};
APH_HD Rng rngInit(uint32_t sampleIndex, uint32_t width, uint32_t px, uint32_t py)
{
    return Rng{py * width + px, sampleIndex};
}
APH_HD void pcg2d(Rng& s)
{
    uint32_t vx = s.x, vy = s.y;
    vx += vy * 1664525u;
    vy += vx * 1664525u;
    vx ^= vx >> 16u;
    vy ^= vy >> 16u;
    vx += vy * 1664525u;
    vy += vx * 1664525u;
    vx ^= vx >> 16u;
    vy ^= vy >> 16u;
    s.x = vx;
    s.y = vy;
}
struct f2 {
    float x, y;
};
// u32 / float(0xFFFFFFFF): both operands are rounded to float, so 1.0 is reachable.
APH_HD f2 random2D(Rng& s)
{
    pcg2d(s);
    const float denom = 4294967296.0f;
    return f2{(float)s.x / denom, (float)s.y / denom};
}

// ---------------- Sampling.slang ----------------
APH_HD f3 sampleUniformDirection(f2 r, float& pdf)  // :20-34
{
    float cosTheta = 1.0f - 2.0f * r.y;
    float sinTheta = sqrtf(1.0f - cosTheta * cosTheta);
    float phi = 2.0f * kPi * r.x;
    pdf = 0.25f * kPiInv;
    return f3{sinTheta * cosf(phi), sinTheta * sinf(phi), cosTheta};
}
APH_HD f3 sampleHemisphere(f2 r, f3 n, float& pdf)  // :39-52
{
    f3 d = sampleUniformDirection(r, pdf);
    d = normalize(n + d);
    pdf = dot(n, d) * kPiInv;
    return d;
}
APH_HD void buildBasis(f3 n, f3& t, f3& b)  // Globals.slang:205-210
{
    f3 bt = fabsf(n.y) < 0.999f ? f3{0.0f, 1.0f, 0.0f} : f3{1.0f, 0.0f, 0.0f};
    t = normalize(cross(bt, n));
    b = cross(n, t);
}
APH_HD f3 sampleCone(f2 r, f3 dir, float cosThetaMax)  // :72-86
{
    f3 xb, yb;
    buildBasis(dir, xb, yb);
    float cosTheta = (1.0f - r.x) + r.x * cosThetaMax;
    float sinTheta = sqrtf(1.0f - cosTheta * cosTheta);
    float phi = r.y * 2.0f * kPi;
    return xb * (cosf(phi) * sinTheta) + yb * (sinf(phi) * sinTheta) + dir * cosTheta;
}
APH_HD float misWeight(float f, float g) { return (f * f) / (f * f + g * g); }  // :101-107

// ---------------- Environment.slang ----------------
APH_HD f2 dirToLatLong(f3 d)  // :35-42
{
    return f2{atan2f(d.x, -d.z) * kPiInv * 0.5f + 0.5f, -asinf(d.y) * kPiInv + 0.5f};
}
APH_HD f3 latLongToDir(f2 uv)  // :46-59
{
    float phi = (uv.x + 0.25f) * 2.0f * kPi;
    float theta = uv.y * kPi;
    float sp = sinf(phi), cp = cosf(phi), st = sinf(theta), ct = cosf(theta);
    return f3{cp * st, ct, sp * st};
}
APH_HD f3 xf4Dir(const float* m, f3 v)
{
    return f3{m[0] * v.x + m[1] * v.y + m[2] * v.z, m[4] * v.x + m[5] * v.y + m[6] * v.z,
              m[8] * v.x + m[9] * v.y + m[10] * v.z};
}
__device__ __forceinline__ f3 texel(const float4* tex, uint32_t w, int x, int y)
{
    float4 t = tex[(size_t)y * w + x];
    return f3{t.x, t.y, t.z};
}
// Repeat addressing; filter 0 = nearest (HGI default sampler), 1 = bilinear.
// A non-finite coordinate reads as 0 (a NaN direction comes from sampleHemisphere
// normalizing n + d with d = -n when a draw is exactly 1.0); without this the
// float-to-int conversion below would be undefined.
__device__ __forceinline__ f3 texLookup(const float4* tex, uint32_t w, uint32_t h, f2 uv, int filter)
{
    if (!isfinite(uv.x)) uv.x = 0.0f;
    if (!isfinite(uv.y)) uv.y = 0.0f;
    if (filter == 0) {
        float fu = uv.x - floorf(uv.x), fv = uv.y - floorf(uv.y);
        int ix = (int)(fu * (float)w), iy = (int)(fv * (float)h);
        ix = ix >= (int)w ? (int)w - 1 : ix;
        iy = iy >= (int)h ? (int)h - 1 : iy;
        return texel(tex, w, ix, iy);
    }
    float x = uv.x * (float)w - 0.5f, y = uv.y * (float)h - 0.5f;
    float x0f = floorf(x), y0f = floorf(y);
    float fx = x - x0f, fy = y - y0f;
    int x0 = (int)x0f, y0 = (int)y0f;
    int xa = x0 % (int)w; if (xa < 0) xa += (int)w;
    int xb = (x0 + 1) % (int)w; if (xb < 0) xb += (int)w;
    int ya = y0 % (int)h; if (ya < 0) ya += (int)h;
    int yb = (y0 + 1) % (int)h; if (yb < 0) yb += (int)h;
    f3 c00 = texel(tex, w, xa, ya), c10 = texel(tex, w, xb, ya);
    f3 c01 = texel(tex, w, xa, yb), c11 = texel(tex, w, xb, yb);
    return lerp(lerp(c00, c10, fx), lerp(c01, c11, fx), fy);
}
// evaluateEnvironment, Environment.slang:98-133
__device__ __forceinline__ f3 evalEnvironment(const DevEnv& e, f3 dir, bool asBackground, uint32_t px, uint32_t py,
                                              uint32_t W, uint32_t H)
{
    if (asBackground) {
        f2 uv;
        if (e.bgUseScreen) uv = f2{(float)px / (float)W, (float)py / (float)H};
        else uv = dirToLatLong(normalize(xf4Dir(e.bgXf, dir)));
        if (e.hasBgTex) return texLookup(e.bgTex, e.bw, e.bh, uv, e.filter);
        return lerp(e.bgTop, e.bgBottom, uv.y);
    }
    f2 uv = dirToLatLong(normalize(xf4Dir(e.lightXf, dir)));
    if (e.hasLightTex) return texLookup(e.lightTex, e.lw, e.lh, uv, e.filter);
    return lerp(e.lightTop, e.lightBottom, uv.y);
}
// sampleEnvironment, Environment.slang:137-186
__device__ __forceinline__ f3 sampleEnvironment(const DevEnv& e, f2 r, f3& L, float& pdf)
{
    if (e.hasLightTex) {
        uint32_t count = e.lw * e.lh;
        uint32_t index = (uint32_t)((float)count * r.x);
        index = index > count - 1 ? count - 1 : index;
        // the alias table's two-way choice, with the chosen texel's pdf and radiance
        // from the same record
        const float4* rec = reinterpret_cast<const float4*>(e.samples + index);
        const float4 a = rec[0], b = rec[1], c = rec[2];
        const bool useAlias = a.y < r.y;
        index = useAlias ? __float_as_uint(a.x) : index;
        pdf = useAlias ? a.w : a.z;
        uint32_t xi = index % e.lw, yi = index / e.lw;
        f3 d = latLongToDir(f2{((float)xi + 0.5f) / (float)e.lw, ((float)yi + 0.5f) / (float)e.lh});
        L = normalize(xf4Dir(e.lightXfInv, d));
        return useAlias ? f3{b.w, c.x, c.y} : f3{b.x, b.y, b.z};
    }
    L = sampleUniformDirection(r, pdf);
    L = normalize(xf4Dir(e.lightXfInv, L));
    f2 uv = dirToLatLong(L);
    return lerp(e.lightTop, e.lightBottom, uv.y);
}

// ---------------- material textures, DefaultMaterial.slang:84-127 ----------------
// applyUVTransform (:97-104; rotateUV :86-93): about the pivot, scale, then
// rotate by the angle in degrees, then subtract the offset.  xf = offset.xy,
// scale.xy, pivot.xy, rotation.
APH_HD f2 applyUVTransform(f2 uv, const float* xf)
{
    float sx = (uv.x - xf[4]) / xf[2], sy = (uv.y - xf[5]) / xf[3];
    float r = xf[6] * (kPi / 180.0f);
    float sa = sinf(r), ca = cosf(r);
    float rx = ca * sx + sa * sy, ry = -sa * sx + ca * sy;
    return f2{(rx - xf[0]) + xf[4], (ry - xf[1]) + xf[5]};
}

// Address mode of an integer texel coordinate (Vulkan rules); border flags a
// coordinate outside the image under APH_ADDRESS_BORDER.
APH_HD int addressTexel(int i, int n, int mode, bool& border)
{
    if (mode == kAddrMirror) {
        int t = i % (2 * n);
        if (t < 0) t += 2 * n;
        return t < n ? t : 2 * n - 1 - t;
    }
    if (mode == kAddrClamp) return i < 0 ? 0 : (i >= n ? n - 1 : i);
    if (mode == kAddrMirrorOnce) {
        int j = i < 0 ? -1 - i : i;
        return j >= n ? n - 1 : j;
    }
    if (mode == kAddrBorder) {
        if (i < 0 || i >= n) border = true;
        return i < 0 ? 0 : (i >= n ? n - 1 : i);
    }
    int t = i % n;  // wrap
    return t < 0 ? t + n : t;
}

// Bilinear sample (one mip level) of a material texture.  Non-finite
// coordinates read as 0 and texel coordinates are clamped to +-2^30 before the
// conversion to int, on both sides of the parity check.
__device__ __forceinline__ float4 sampleTexture(const DevTexture& t, f2 uv)
{
    if (!isfinite(uv.x)) uv.x = 0.0f;
    if (!isfinite(uv.y)) uv.y = 0.0f;
    float x = uv.x * (float)t.w - 0.5f, y = uv.y * (float)t.h - 0.5f;
    float x0f = floorf(x), y0f = floorf(y);
    float fx = x - x0f, fy = y - y0f;
    x0f = fminf(fmaxf(x0f, -1073741824.0f), 1073741824.0f);
    y0f = fminf(fmaxf(y0f, -1073741824.0f), 1073741824.0f);
    int x0 = (int)x0f, y0 = (int)y0f;
    float4 c[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        bool border = false;
        int xi = addressTexel(x0 + (k & 1), (int)t.w, t.addrU, border);
        int yi = addressTexel(y0 + (k >> 1), (int)t.h, t.addrV, border);
        c[k] = border ? make_float4(1.0f, 1.0f, 1.0f, 1.0f) : t.texels[(size_t)yi * t.w + xi];
    }
    auto mix = [](float4 a, float4 b, float w) {
        return make_float4(a.x + (b.x - a.x) * w, a.y + (b.y - a.y) * w, a.z + (b.z - a.z) * w, a.w + (b.w - a.w) * w);
    };
    return mix(mix(c[0], c[1], fx), mix(c[2], c[3], fx), fy);
}

// calculateNormalFromMap (:107-127), tangent space: the texel remapped to
// [-1, 1] in the frame (T, normalize(N x T), N).
APH_HD f3 normalFromMap(f3 texel, f3 N, f3 T)
{
    f3 v = texel * 2.0f - splat(1.0f);
    f3 B = normalize(cross(N, T));
    return normalize(T * v.x + B * v.y + N * v.z);
}

// ---------------- shading data, Geometry.slang:131-265 ----------------
struct Shading {
    f3 normal, position, geomPosition, tangent, bitangent;
    // the hit material's base colour and specular roughness after its textures
    // (the BSDF reads these, not the material's constants)
    f3 baseColor;
    float specularRoughness;
};

__device__ __forceinline__ f3 ld3(const float* p) { return f3{p[0], p[1], p[2]}; }

__device__ __forceinline__ void computeShading(const InstShade& in, uint32_t prim, float bu, float bv, Shading& sh)
{
    float b[3] = {1.0f - bu - bv, bu, bv};
    uint32_t i0 = in.idx[3 * prim + 0], i1 = in.idx[3 * prim + 1], i2 = in.idx[3 * prim + 2];
    f3 p[3] = {ld3(in.pos + 3 * (size_t)i0), ld3(in.pos + 3 * (size_t)i1), ld3(in.pos + 3 * (size_t)i2)};
    f3 geom = p[0] * b[0] + p[1] * b[1] + p[2] * b[2];
    f3 objN, objP;
    if (in.nrm) {
        f3 n[3] = {ld3(in.nrm + 3 * (size_t)i0), ld3(in.nrm + 3 * (size_t)i1), ld3(in.nrm + 3 * (size_t)i2)};
        f3 nrm = n[0] * b[0] + n[1] * b[1] + n[2] * b[2];
        objN = normalize(nrm);
        // computeShadingPosition (:131-148): project onto the vertex tangent planes.
        f3 sp = f3{0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int i = 0; i < 3; i++) sp = sp + (geom - n[i] * dot(geom - p[i], n[i])) * b[i];
        objP = dot(sp - geom, nrm) > 0.0f ? sp : geom;
    } else {
        objN = normalize(cross(p[1] - p[0], p[2] - p[0]));
        objP = geom;
    }
    sh.position = xfPoint(in.o2w, objP);
    sh.geomPosition = xfPoint(in.o2w, geom);
    sh.normal = normalize(xfVector(in.o2w, objN));
    f3 up = fabsf(sh.normal.y) < 0.999f ? f3{0.0f, 1.0f, 0.0f} : f3{1.0f, 0.0f, 0.0f};
    sh.tangent = normalize(cross(sh.normal, up));
    sh.bitangent = cross(sh.normal, sh.tangent);
}

// ---------------- Standard Surface, StandardSurfaceBSDF.slang ----------------
APH_HD bool sameHemi(f3 a, f3 b, f3 n) { return dot(a, n) * dot(b, n) > 0.0f; }
APH_HD f2 roughAniso(float roughness, float anisotropy)  // :31-56
{
    float rs = clampf(roughness * roughness, kFloatEps, 1.0f);
    if (anisotropy > 0.0f) {
        float aspect = sqrtf(1.0f - clampf(anisotropy, 0.0f, 0.98f));
        return f2{fminf(rs / aspect, 1.0f), rs * aspect};
    }
    return f2{rs, rs};
}
APH_HD float pow5(float x) { return sqr(sqr(x)) * x; }
APH_HD f3 mxMix(f3 bg, f3 fg, float w) { return lerp(bg, fg, clampf(w, 0.0f, 1.0f)); }
APH_HD float orenNayar(f3 L, f3 V, f3 N, float NdotL, float roughness)  // :76-89
{
    float LdotV = dot(L, V), NdotV = fabsf(dot(N, V));
    float t = LdotV - NdotL * NdotV;
    t = t > 0.0f ? t / fmaxf(NdotL, NdotV) : 0.0f;
    float s2 = sqr(roughness * kPi);
    float A = 1.0f - 0.5f * (s2 / (s2 + 0.33f));
    float B = 0.45f * s2 / (s2 + 0.09f);
    return A + B * t;
}
APH_HD f3 diffuseBrdf(f3 L, f3 V, float weight, f3 color, float roughness, f3 N)  // :91-112
{
    if (weight < kFloatEps || !sameHemi(L, V, N)) return splat(0.0f);
    float NdotL = fmaxf(dot(L, N), 0.0f);
    f3 r = color * (weight * NdotL * kPiInv);
    if (roughness > 0.0f) r = r * orenNayar(L, V, N, NdotL, roughness);
    return r;
}
APH_HD float avgRough(f2 r) { return sqrtf(r.x * r.y); }
APH_HD float iorToF0(float ior) { return sqr((ior - 1.0f) / (ior + 1.0f)); }
APH_HD float ggxNdfMx(f3 X, f3 Y, f3 H, float NdotH, float ax, float ay)  // :128-134
{
    float d = sqr(dot(X, H) / ax) + sqr(dot(Y, H) / ay) + sqr(NdotH);
    return 1.0f / (kPi * ax * ay * sqr(d));
}
APH_HD float fresnelSchlick(float cosTheta, float F0)  // :137-142
{
    float x = clampf(1.0f - cosTheta, 0.0f, 1.0f);
    return F0 + (1.0f - F0) * pow5(x);
}
APH_HD float smithG(float NdotL, float NdotV, float alpha)  // :146-152
{
    float a2 = sqr(alpha);
    float lL = sqrtf(a2 + (1.0f - a2) * sqr(NdotL));
    float lV = sqrtf(a2 + (1.0f - a2) * sqr(NdotV));
    return 2.0f / (lL / NdotL + lV / NdotV);
}
APH_HD float dirAlbedoFit(float NdotV, float roughness, float F0, float F90)  // :158-171
{
    float r0 = roughness * -1.0f + 1.0f, r1 = roughness * -0.0275f + 0.0425f;
    float r2 = roughness * -0.572f + 1.04f, r3 = roughness * 0.022f + -0.04f;
    float a004 = fminf(r0 * r0, exp2f(-9.28f * NdotV)) * r0 + r1;
    return F0 * (-1.04f * a004 + r2) + F90 * (1.04f * a004 + r3);
}
APH_HD float energyComp(float NdotV, float roughness, float Fss)  // :175-184
{
    float Ess = dirAlbedoFit(NdotV, roughness, 1.0f, 1.0f);
    return 1.0f + Fss * (1.0f - Ess) / Ess;
}
APH_HD f3 energyComp3(float NdotV, float roughness, f3 Fss)
{
    float Ess = dirAlbedoFit(NdotV, roughness, 1.0f, 1.0f);
    return splat(1.0f) + Fss * (1.0f - Ess) / Ess;
}
APH_HD f3 dielectricBrdf(f3 L, f3 V, float weight, f3 tint, float ior, f2 rough, f3 N, f3 X, f3 base)  // :186-220
{
    if (weight < kFloatEps) return base;
    float NdotL = fabsf(dot(N, L)), NdotV = fabsf(dot(N, V));
    f3 Y = normalize(cross(N, X));
    f3 H = normalize(L + V);
    float NdotH = dot(N, H), VdotH = dot(V, H);
    float ar = avgRough(rough);
    float F0 = iorToF0(ior);
    float D = ggxNdfMx(X, Y, H, NdotH, rough.x, rough.y);
    float F = fresnelSchlick(VdotH, F0);
    float G = smithG(NdotL, NdotV, ar);
    float comp = energyComp(NdotV, ar, F);
    float dirAlbedo = dirAlbedoFit(NdotV, ar, F0, 1.0f) * comp;
    return fabs3(tint * (D * F * G * comp * weight) / (4.0f * NdotV)) + base * (1.0f - dirAlbedo * weight);
}
APH_HD void artisticToComplexIor(f3 refl, f3 edge, f3& n, f3& k)  // :222-236
{
    f3 r = f3{clampf(refl.x, 0.0f, 0.99f), clampf(refl.y, 0.0f, 0.99f), clampf(refl.z, 0.0f, 0.99f)};
    f3 rs = f3{sqrtf(r.x), sqrtf(r.y), sqrtf(r.z)};
    f3 nmin = (splat(1.0f) - r) / (splat(1.0f) + r);
    f3 nmax = (splat(1.0f) + rs) / (splat(1.0f) - rs);
    n = lerp(nmax, nmin, edge);
    f3 np1 = n + splat(1.0f), nm1 = n - splat(1.0f);
    f3 k2 = (np1 * np1 * r - nm1 * nm1) / (splat(1.0f) - r);
    k = f3{sqrtf(fmaxf(k2.x, 0.0f)), sqrtf(fmaxf(k2.y, 0.0f)), sqrtf(fmaxf(k2.z, 0.0f))};
}
APH_HD f3 fresnelConductor(float c, f3 n, f3 k)  // :238-250
{
    float c2 = c * c;
    f3 n2k2 = n * n + k * k;
    f3 nc2 = n * (2.0f * c);
    f3 rsa = n2k2 + splat(c2), rpa = n2k2 * c2 + splat(1.0f);
    f3 rs = (rsa - nc2) / (rsa + nc2), rp = (rpa - nc2) / (rpa + nc2);
    return (rs + rp) * 0.5f;
}
APH_HD f3 conductorBrdf(f3 L, f3 V, float weight, f3 refl, f3 edge, f2 rough, f3 N, f3 X)  // :252-285
{
    if (weight < kFloatEps || !sameHemi(L, V, N)) return splat(0.0f);
    float NdotL = fabsf(dot(N, L)), NdotV = fabsf(dot(N, V));
    f3 Y = normalize(cross(N, X));
    f3 H = normalize(L + V);
    float NdotH = dot(N, H), VdotH = dot(V, H);
    f3 n, k;
    artisticToComplexIor(refl, edge, n, k);
    float ar = avgRough(rough);
    float D = ggxNdfMx(X, Y, H, NdotH, rough.x, rough.y);
    f3 F = fresnelConductor(VdotH, n, k);
    float G = smithG(NdotL, NdotV, ar);
    f3 comp = energyComp3(NdotV, ar, F);
    return F * comp * (D * G * weight) / (4.0f * NdotV);
}
APH_HD float sheenNdf(float cosTheta, float roughness)  // :292-299
{
    float inv = 1.0f / roughness;
    float sin2 = 1.0f - cosTheta * cosTheta;
    return (2.0f + inv) * powf(sin2, inv * 0.5f) / (2.0f * kPi);
}
// Sheen directional-albedo LUT (16x16), the MaterialX table of :307-337.
__device__ __constant__ const float kSheenLut[256] = {1.6177f,
    0.978927f, 0.618938f, 0.391714f, 0.245177f, 0.150234f, 0.0893475f, 0.0511377f, 0.0280191f, 0.0144204f,
    0.00687674f, 0.00295935f, 0.00111049f, 0.000336768f, 7.07119e-05f, 6.22646e-06f, 1.1084f, 0.813928f,
    0.621389f, 0.479304f, 0.370299f, 0.284835f, 0.21724f, 0.163558f, 0.121254f, 0.0878921f, 0.0619052f,
    0.0419894f, 0.0270556f, 0.0161443f, 0.00848212f, 0.00342323f, 0.930468f, 0.725652f, 0.586532f, 0.479542f,
    0.393596f, 0.322736f, 0.26353f, 0.213565f, 0.171456f, 0.135718f, 0.105481f, 0.0800472f, 0.0588117f,
    0.0412172f, 0.0268329f, 0.0152799f, 0.833791f, 0.671201f, 0.558957f, 0.471006f, 0.398823f, 0.337883f,
    0.285615f, 0.240206f, 0.200696f, 0.16597f, 0.135422f, 0.10859f, 0.0850611f, 0.0644477f, 0.0464763f,
    0.0308878f, 0.771692f, 0.633819f, 0.537877f, 0.461939f, 0.398865f, 0.344892f, 0.297895f, 0.256371f,
    0.219562f, 0.186548f, 0.156842f, 0.130095f, 0.10598f, 0.0841919f, 0.0645311f, 0.04679f, 0.727979f,
    0.606373f, 0.52141f, 0.453769f, 0.397174f, 0.348337f, 0.305403f, 0.267056f, 0.232655f, 0.201398f,
    0.17286f, 0.146756f, 0.122808f, 0.100751f, 0.0804254f, 0.0616485f, 0.695353f, 0.585281f, 0.508227f,
    0.44667f, 0.394925f, 0.350027f, 0.310302f, 0.274561f, 0.242236f, 0.212604f, 0.185281f, 0.16002f, 0.13657f,
    0.114693f, 0.0942543f, 0.0750799f, 0.669981f, 0.568519f, 0.497442f, 0.440542f, 0.392567f, 0.350786f,
    0.313656f, 0.280075f, 0.249533f, 0.221359f, 0.195196f, 0.170824f, 0.148012f, 0.126537f, 0.106279f,
    0.0870713f, 0.649644f, 0.554855f, 0.488453f, 0.435237f, 0.390279f, 0.351028f, 0.316036f, 0.284274f,
    0.255266f, 0.228387f, 0.203297f, 0.179796f, 0.157665f, 0.136695f, 0.116774f, 0.0977403f, 0.632951f,
    0.543489f, 0.480849f, 0.430619f, 0.388132f, 0.350974f, 0.317777f, 0.287562f, 0.259885f, 0.234153f,
    0.210041f, 0.187365f, 0.165914f, 0.145488f, 0.125983f, 0.10724f, 0.61899f, 0.533877f, 0.47433f, 0.426573f,
    0.386145f, 0.35075f, 0.319078f, 0.290197f, 0.263681f, 0.238971f, 0.215746f, 0.193838f, 0.173043f,
    0.153167f, 0.134113f, 0.115722f, 0.607131f, 0.52564f, 0.468678f, 0.423001f, 0.38432f, 0.35043f, 0.320072f,
    0.292349f, 0.266856f, 0.243055f, 0.220636f, 0.199438f, 0.179264f, 0.159926f, 0.141332f, 0.123323f,
    0.596927f, 0.518497f, 0.463731f, 0.419829f, 0.382647f, 0.350056f, 0.320842f, 0.294137f, 0.269549f,
    0.246564f, 0.224875f, 0.204331f, 0.18474f, 0.165919f, 0.147778f, 0.130162f, 0.588052f, 0.512241f,
    0.459365f, 0.416996f, 0.381114f, 0.349657f, 0.321448f, 0.295641f, 0.271862f, 0.24961f, 0.228584f,
    0.208643f, 0.189596f, 0.171266f, 0.153566f, 0.136341f, 0.580257f, 0.506717f, 0.455481f, 0.41445f,
    0.379708f, 0.34925f, 0.321929f, 0.296923f, 0.273869f, 0.252279f, 0.231859f, 0.212472f, 0.193933f,
    0.176066f, 0.158788f, 0.141945f, 0.573355f, 0.5018f, 0.452005f, 0.412151f, 0.378416f, 0.348844f,
    0.322316f, 0.298028f, 0.275627f, 0.254638f, 0.234772f, 0.215896f, 0.197828f, 0.180398f, 0.163522f,
    0.147049f};
__device__ __forceinline__ float sheenAlbedo(float cosTheta, float roughness)  // :341-362
{
    float x = cosTheta * 15.0f, y = roughness * 15.0f;
    // NaN reads row/column 0 and indices stay inside the table; the callers'
    // finite arguments lie in [0, 1] and are unaffected.
    if (!(x >= 0.0f)) x = 0.0f;
    if (!(y >= 0.0f)) y = 0.0f;
    int ix = min((int)fminf(x, 15.0f), 15), iy = min((int)fminf(y, 15.0f), 15);
    int ix2 = min(max(ix + 1, 0), 15), iy2 = min(max(iy + 1, 0), 15);
    float fx = x - (float)ix, fy = y - (float)iy;
    float v1 = lerp(kSheenLut[iy * 16 + ix], kSheenLut[iy * 16 + ix2], fx);
    float v2 = lerp(kSheenLut[iy2 * 16 + ix], kSheenLut[iy2 * 16 + ix2], fx);
    return clampf(lerp(v1, v2, fy), 0.0f, 1.0f);
}
__device__ __forceinline__ f3 sheenBrdf(f3 L, f3 V, float weight, f3 color, float roughness, f3 N, f3 base)
{  // :366-410
    if (weight < kFloatEps) return base;
    if (dot(N, V) < 0.0f) N = -N;
    f3 H = normalize(L + V);
    float NdotL = clampf(dot(N, L), kFloatEps, 1.0f), NdotV = clampf(dot(N, V), kFloatEps, 1.0f);
    float NdotH = clampf(dot(N, H), kFloatEps, 1.0f);
    float alpha = clampf(roughness, kFloatEps, 1.0f);
    float D = sheenNdf(NdotH, alpha);
    f3 fr = color * D / (4.0f * (NdotL + NdotV - NdotL * NdotV));
    float dirAlbedo = sheenAlbedo(NdotV, alpha);
    return fr * (NdotL * weight) + base * (1.0f - dirAlbedo * weight);
}
__device__ __forceinline__ f3 sheenSpecular(f3 L, f3 V, float weight, f3 color, float roughness, f3 N)
{  // :414-458
    if (!sameHemi(L, V, N)) return splat(0.0f);
    if (dot(N, V) < 0.0f) N = -N;
    f3 H = normalize(L + V);
    float NdotL = clampf(dot(N, L), kFloatEps, 1.0f), NdotV = clampf(dot(N, V), kFloatEps, 1.0f);
    float NdotH = clampf(dot(N, H), kFloatEps, 1.0f);
    float alpha = clampf(roughness, kFloatEps, 1.0f);
    float D = sheenNdf(NdotH, alpha);
    f3 fr = color * D / (4.0f * (NdotL + NdotV - NdotL * NdotV));
    return fr * (NdotL * weight);
}
APH_HD f3 subsurfaceBrdf(f3 L, float weight, f3 color, f3 N)  // :462-473
{
    float NdotL = dot(L, -N);
    if (NdotL <= 0.0f || weight < kFloatEps) return splat(0.0f);
    return color * (weight * NdotL * kPiInv);
}
APH_HD f3 rotateVector(f3 v, float amountDeg, f3 axis)  // mx_rotate_vector3, :477-502
{
    float a = amountDeg * (kPi / 180.0f);
    axis = normalize(axis);
    float s = sinf(a), c = cosf(a), oc = 1.0f - c;
    return f3{(oc * axis.x * axis.x + c) * v.x + (oc * axis.x * axis.y - axis.z * s) * v.y +
                  (oc * axis.z * axis.x + axis.y * s) * v.z,
              (oc * axis.x * axis.y + axis.z * s) * v.x + (oc * axis.y * axis.y + c) * v.y +
                  (oc * axis.y * axis.z - axis.x * s) * v.z,
              (oc * axis.z * axis.x - axis.y * s) * v.x + (oc * axis.y * axis.z + axis.x * s) * v.y +
                  (oc * axis.z * axis.z + c) * v.z};
}
APH_HD f3 dielectricDiffuse(f3 L, f3 V, float weight, float ior, f2 rough, f3 N, f3 albedo)  // :506-540
{
    if (weight < kFloatEps || !sameHemi(L, V, N)) return splat(0.0f);
    float NdotV = fabsf(dot(N, V));
    f3 H = normalize(L + V);
    float VdotH = dot(V, H);
    float ar = avgRough(rough);
    float F0 = iorToF0(ior);
    float F = fresnelSchlick(VdotH, F0);
    float comp = energyComp(NdotV, ar, F);
    float dirAlbedo = dirAlbedoFit(NdotV, ar, F0, 1.0f) * comp;
    return albedo * (1.0f - dirAlbedo) * (weight * kPiInv);
}
APH_HD f3 dielectricSpecular(f3 L, f3 V, f3 tint, float ior, f2 rough, f3 N, f3 X)  // :545-584
{
    if (!sameHemi(L, V, N)) return splat(0.0f);
    float NdotL = fabsf(dot(N, L)), NdotV = fabsf(dot(N, V));
    f3 Y = normalize(cross(N, X));
    f3 H = normalize(L + V);
    float NdotH = dot(N, H), VdotH = dot(V, H);
    float ar = avgRough(rough);
    float F0 = iorToF0(ior);
    float D = ggxNdfMx(X, Y, H, NdotH, rough.x, rough.y);
    float F = fresnelSchlick(VdotH, F0);
    float G = smithG(NdotL, NdotV, ar);
    float comp = energyComp(NdotV, ar, F);
    return fabs3(tint * (D * F * G * comp) / (4.0f * NdotV * NdotL));
}
// BSDFCommon.slang helpers
APH_HD float ggxNdf(f3 H, float ax, float ay)  // :245-250
{
    float e = sqr(H.x) / sqr(ax) + sqr(H.y) / sqr(ay) + sqr(H.z);
    return 1.0f / (kPi * ax * ay * sqr(e));
}
APH_HD float smithLambda(f3 v, float ax, float ay)  // :208-213
{
    float s = (sqr(ax) * sqr(v.x) + sqr(ay) * sqr(v.y)) / sqr(v.z);
    return (-1.0f + sqrtf(1.0f + s)) * 0.5f;
}
APH_HD float ggxVndf(f3 V, f3 H, float ax, float ay)  // :255-269
{
    float VdotH = dot(V, H);
    if (VdotH <= 0.0f) return 0.0f;
    float G1 = dot(V, H) <= 0.0f ? 0.0f : 1.0f / (1.0f + smithLambda(V, ax, ay));
    return ggxNdf(H, ax, ay) * G1 * VdotH / fabsf(V.z);
}
APH_HD f3 sampleVndf(f3 V, float ax, float ay, f2 r)  // :274-299 (Heitz 2018)
{
    f3 Vh = normalize(f3{ax * V.x, ay * V.y, V.z});
    float lensq = Vh.x * Vh.x + Vh.y * Vh.y;
    f3 T1 = lensq > 0.0f ? f3{-Vh.y, Vh.x, 0.0f} / sqrtf(lensq) : f3{1.0f, 0.0f, 0.0f};
    f3 T2 = cross(Vh, T1);
    float rr = sqrtf(r.x), phi = 2.0f * kPi * r.y;
    float t1 = rr * cosf(phi), t2 = rr * sinf(phi);
    float s = 0.5f * (1.0f + Vh.z);
    t2 = (1.0f - s) * sqrtf(1.0f - t1 * t1) + s * t2;
    f3 Nh = T1 * t1 + T2 * t2 + Vh * sqrtf(fmaxf(0.0f, 1.0f - t1 * t1 - t2 * t2));
    return normalize(f3{ax * Nh.x, ay * Nh.y, fmaxf(0.0f, Nh.z)});
}
APH_HD f3 computeRefraction(f3 V, float ior, bool& reflected)  // BSDFCommon.slang:122-152
{
    reflected = false;
    bool entering = V.z > 0.0f;
    f3 N = f3{0.0f, 0.0f, entering ? 1.0f : -1.0f};
    float eta = entering ? 1.0f / ior : ior / 1.0f;
    f3 L = refract(-V, N, eta);
    if (length(L) == 0.0f) {
        L = f3{-V.x, -V.y, V.z};
        reflected = true;
    }
    return L;
}

APH_HD f3 toTangent(f3 i, const Shading& s) { return f3{dot(i, s.tangent), dot(i, s.bitangent), dot(i, s.normal)}; }
APH_HD f3 toWorld(f3 i, const Shading& s) { return s.tangent * i.x + s.bitangent * i.y + s.normal * i.z; }

// Quantities shared by evaluate / sample (common prologue of :598-629, :862-905, :1093-1143).
// A gamma of 1 and a rotation of 0 are identities (pow(x, 1) = x; a rotation by
// sin 0 = 0, cos 0 = 1 returns v up to the sign of zero components), so the
// common material skips the powf and sincos work the general case needs.
struct Prologue {
    f2 mainRough;
    f3 mainTangent, coatTangent, diffuseColor;
    float coatGamma;
};
__device__ __forceinline__ f3 coatGammaPow(f3 c, float gamma) { return gamma == 1.0f ? c : pow3(c, gamma); }
__device__ __forceinline__ f3 tangentRotation(const Shading& sh, float turns)
{
    return turns == 0.0f ? sh.tangent : rotateVector(sh.tangent, turns * 360.0f, sh.normal);
}
__device__ __forceinline__ Prologue prologue(const DevMaterial& m, const Shading& sh)
{
    Prologue p;
    float caf = m.coatAffectRoughness * m.coat * m.coatRoughness;
    float car = lerp(sh.specularRoughness, 1.0f, caf);
    p.coatGamma = clampf(m.coat, 0.0f, 1.0f) * m.coatAffectColor + 1.0f;
    p.mainRough = roughAniso(car, m.specularAnisotropy);
    p.mainTangent = tangentRotation(sh, m.specularRotation);
    p.coatTangent = tangentRotation(sh, m.coatRotation);
    p.diffuseColor = coatGammaPow(sh.baseColor, p.coatGamma);
    return p;
}

// evaluateMaterial, :598-718 (includes the cosine term).
__device__ f3 evalMaterial(const DevMaterial& m, const Shading& sh, f3 Vw, f3 Lw)
{
    Prologue p = prologue(m, sh);
    f3 N = sh.normal;
    f3 diffuse = diffuseBrdf(Lw, Vw, m.base, p.diffuseColor, m.diffuseRoughness, N);
    f3 ssMix = diffuse;
    if (m.subsurface > 0.0f)
        ssMix = mxMix(diffuse, subsurfaceBrdf(Lw, 1.0f, coatGammaPow(m.subsurfaceColor, p.coatGamma), N),
                      m.subsurface);
    if (dot(N, Vw) * dot(N, Lw) < 0.0f) return ssMix;
    f3 sheenDiff = ssMix;
    if (m.sheen > 0.0f) sheenDiff = sheenBrdf(Lw, Vw, m.sheen, m.sheenColor, m.sheenRoughness, N, ssMix);
    sheenDiff = sheenDiff * (1.0f - m.transmission);
    f3 spec = sheenDiff;
    if (m.specular > 0.0f)
        spec = dielectricBrdf(Lw, Vw, m.specular, m.specularColor, m.specularIOR, p.mainRough, N, p.mainTangent,
                              sheenDiff);
    f3 metalMix = spec;
    if (m.metalness > 0.0f) {
        f3 metal = conductorBrdf(Lw, Vw, 1.0f, sh.baseColor * m.base, m.specularColor * m.specular, p.mainRough, N,
                                 p.mainTangent);
        metalMix = mxMix(spec, metal, m.metalness);
    }
    f3 fin = metalMix;
    if (m.coat > 0.0f) {
        f2 cr = roughAniso(m.coatRoughness, m.coatAnisotropy);
        f3 att = lerp(splat(1.0f), m.coatColor, m.coat);
        fin = dielectricBrdf(Lw, Vw, m.coat, m.coatColor, m.coatIOR, cr, N, p.coatTangent, att * metalMix);
    }
    return fin;
}

APH_HD float ggxPdf(f3 V, f2 a, f3 L)  // calculateGGXPDF, :737-746
{
    f3 H = normalize(V + L);
    return ggxVndf(V, H, a.x, a.y) / (4.0f * dot(V, H));
}
APH_HD f3 ggxDirPdf(const Shading& sh, f3 V, f2 a, f2 r, f3& Lw, float& pdf)  // :751-773
{
    bool flip = V.z <= 0.0f;
    f3 H = sampleVndf(flip ? -V : V, a.x, a.y, r);
    if (flip) H = -H;
    f3 L = reflect(-V, H);
    pdf = ggxVndf(V, H, a.x, a.y) / (4.0f * dot(V, H));
    Lw = toWorld(L, sh);
    return L;
}

// evaluateMaterialAndPDF, :862-1088 (progressive path: lobe probabilities unadjusted).
// As in sampleMaterial, the lobe is chosen first and the GGX lobes share one pdf
// and (coat, specular) one dielectric evaluation.
__device__ f3 evalMaterialPdf(const DevMaterial& m, const Shading& sh, f3 Vw, f3 Lw, f2 r, float& pdf, int& lobe)
{
    f3 N = sh.normal;
    f3 V = toTangent(Vw, sh), L = toTangent(Lw, sh);
    float NdotV = fabsf(V.z);
    pdf = 0.0f;
    lobe = kNoSecondaryRay;
    Prologue p = prologue(m, sh);
    f3 res = splat(0.0f);
    float coatLum = luminance(m.coatColor * (fresnelSchlick(NdotV, iorToF0(m.coatIOR)) * m.coat));
    float specLum = 0.0f, sheenAlpha = 0.0f;
    f2 ga = f2{0.0f, 0.0f};
    if (r.x < coatLum) {
        r.x = r.x / coatLum;
        ga = roughAniso(m.coatRoughness, m.coatAnisotropy);
        lobe = kCoatLobe;
    } else {
        r.x = (r.x - coatLum) / (1.0f - coatLum);
        f2 sa = roughAniso(sh.specularRoughness, m.specularAnisotropy);
        if (r.x < m.metalness) {
            r.x = r.x / m.metalness;
            ga = sa;
            lobe = kMetalLobe;
        } else {
            r.x = (r.x - m.metalness) / (1.0f - m.metalness);
            specLum = luminance(m.specularColor * (fresnelSchlick(NdotV, iorToF0(m.specularIOR)) * m.specular));
            if (r.x < specLum) {
                r.x = r.x / specLum;
                ga = sa;
                lobe = kGlossyLobe;
            } else {
                r.x = (r.x - specLum) / (1.0f - specLum);
                sheenAlpha = clampf(m.sheenRoughness, kFloatEps, 1.0f);
                if (r.x < m.sheen) {
                    r.x = r.x / m.sheen;
                    lobe = kSheenLobe;
                } else {
                    r.x = (r.x - m.sheen) / (1.0f - m.sheen);
                    lobe = kDiffuseLobe;
                }
            }
        }
    }
    if (lobe == kCoatLobe || lobe == kMetalLobe || lobe == kGlossyLobe) {
        pdf = ggxPdf(V, ga, L);
        if (lobe == kMetalLobe) {
            res = conductorBrdf(Lw, Vw, 1.0f, sh.baseColor * m.base, m.specularColor * m.specular, p.mainRough, N,
                                p.mainTangent) /
                  fabsf(L.z);
        } else {
            const bool coat = lobe == kCoatLobe;
            res = dielectricSpecular(Lw, Vw, coat ? m.coatColor * m.coat : m.specularColor * m.specular,
                                     coat ? m.coatIOR : m.specularIOR, coat ? ga : p.mainRough, N,
                                     coat ? p.coatTangent : p.mainTangent);
            if (!coat) res = res / specLum;
        }
    } else if (lobe == kSheenLobe) {
        f3 H = normalize(L + V);
        res = sheenSpecular(Lw, Vw, m.sheen, m.sheenColor, m.sheenRoughness, N);
        pdf = sheenNdf(H.z, sheenAlpha);
    } else {
        pdf = dot(N, Lw) * kPiInv;
        res = dielectricDiffuse(Lw, Vw, m.base, m.specularIOR, p.mainRough, N, p.diffuseColor);
        res = res * (1.0f - sheenAlbedo(NdotV, sheenAlpha) * m.sheen);
    }
    if (lobe != kCoatLobe) res = res * lerp(splat(1.0f), m.coatColor, m.coat);
    return res * fabsf(L.z);
}

// sampleMaterial, :1093-1372.  The lobe is chosen first; the three GGX lobes
// (coat, metal, specular) then share one VNDF sample with their own roughness,
// and coat and specular one dielectric evaluation with their own parameters - the
// same calls on the same arguments as lobe-by-lobe code, but a wave whose lanes
// chose different GGX lobes runs the sampling once instead of up to three times.
__device__ f3 sampleMaterial(const DevMaterial& m, const Shading& sh, f3 Vw, f2 r, f3& Lw, float& pdf, int& lobe)
{
    f3 N = sh.normal;
    f3 V = toTangent(Vw, sh);
    float NdotV = fabsf(V.z);
    f3 L = splat(0.0f);
    Lw = f3{1.0f, 0.0f, 0.0f};
    pdf = 0.0f;
    lobe = kNoSecondaryRay;
    Prologue p = prologue(m, sh);
    f3 res = splat(0.0f);
    float coatLum = luminance(m.coatColor * (fresnelSchlick(NdotV, iorToF0(m.coatIOR)) * m.coat));
    float specLum = 0.0f, sheenAlpha = 0.0f;
    f2 ga = f2{0.0f, 0.0f};  // roughness of the chosen GGX lobe
    if (r.x < coatLum) {
        r.x = r.x / coatLum;
        ga = roughAniso(m.coatRoughness, m.coatAnisotropy);
        lobe = kCoatLobe;
    } else {
        r.x = (r.x - coatLum) / (1.0f - coatLum);
        f2 sa = roughAniso(sh.specularRoughness, m.specularAnisotropy);
        if (r.x < m.metalness) {
            r.x = r.x / m.metalness;
            ga = sa;
            lobe = kMetalLobe;
        } else {
            r.x = (r.x - m.metalness) / (1.0f - m.metalness);
            specLum = luminance(m.specularColor * (fresnelSchlick(NdotV, iorToF0(m.specularIOR)) * m.specular));
            if (r.x < specLum) {
                r.x = r.x / specLum;
                ga = sa;
                lobe = kGlossyLobe;
            } else {
                r.x = (r.x - specLum) / (1.0f - specLum);
                sheenAlpha = clampf(m.sheenRoughness, kFloatEps, 1.0f);
                if (r.x < m.sheen) {
                    r.x = r.x / m.sheen;
                    lobe = kSheenLobe;
                } else {
                    r.x = (r.x - m.sheen) / (1.0f - m.sheen);
                    if (r.x < m.transmission) {
                        r.x = r.x / m.transmission;
                        lobe = kTransmissionLobe;
                    } else {
                        r.x = (r.x - m.transmission) / (1.0f - m.transmission);
                        lobe = kDiffuseLobe;
                    }
                }
            }
        }
    }
    if (lobe == kCoatLobe || lobe == kMetalLobe || lobe == kGlossyLobe) {
        L = ggxDirPdf(sh, V, ga, r, Lw, pdf);
        if (lobe == kMetalLobe) {
            res = conductorBrdf(Lw, Vw, 1.0f, sh.baseColor * m.base, m.specularColor * m.specular, p.mainRough, N,
                                p.mainTangent) /
                  fabsf(L.z);
        } else {
            const bool coat = lobe == kCoatLobe;
            res = dielectricSpecular(Lw, Vw, coat ? m.coatColor * m.coat : m.specularColor * m.specular,
                                     coat ? m.coatIOR : m.specularIOR, coat ? ga : p.mainRough, N,
                                     coat ? p.coatTangent : p.mainTangent) /
                  (coat ? coatLum : specLum);
        }
    } else if (lobe == kSheenLobe) {
        L = sampleHemisphere(r, f3{0.0f, 0.0f, V.z > 0.0f ? 1.0f : -1.0f}, pdf);
        Lw = toWorld(L, sh);
        f3 H = normalize(L + V);
        res = sheenSpecular(Lw, Vw, m.sheen, m.sheenColor, m.sheenRoughness, N);
        pdf = sheenNdf(H.z, sheenAlpha);
    } else if (lobe == kTransmissionLobe) {
        bool reflected = false;
        L = -V;
        if (!m.thinWalled) L = computeRefraction(V, m.specularIOR, reflected);
        Lw = toWorld(L, sh);
        res = m.transmissionColor / fabsf(L.z);
        pdf = 1.0f;
        lobe = reflected ? kGlossyLobe : kTransmissionLobe;
    } else {
        L = sampleHemisphere(r, f3{0.0f, 0.0f, V.z > 0.0f ? 1.0f : -1.0f}, pdf);
        Lw = toWorld(L, sh);
        res = dielectricDiffuse(Lw, Vw, m.base, m.specularIOR, p.mainRough, N, p.diffuseColor);
        res = res * (1.0f - sheenAlbedo(NdotV, sheenAlpha) * m.sheen);
    }
    if (lobe != kCoatLobe) res = res * lerp(splat(1.0f), m.coatColor, m.coat);
    return res * fabsf(L.z);
}

// adjustRadiance, PathTracingCommon.slang:59-82
APH_HD f3 adjustRadiance(float maxLum, bool displayErrors, f3 r)
{
    float lum = luminance(r);
    if (lum > maxLum) r = r * (maxLum / lum);
    if (isnan(r.x) || isnan(r.y) || isnan(r.z)) return displayErrors ? f3{INFINITY, 0.0f, 0.0f} : splat(0.0f);
    if (isinf(r.x) || isinf(r.y) || isinf(r.z)) return displayErrors ? f3{0.0f, INFINITY, 0.0f} : splat(0.0f);
    return r;
}

}  // namespace aph
