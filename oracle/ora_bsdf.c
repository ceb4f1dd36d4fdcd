/* Oracle: Standard Surface BSDF (test infrastructure only).
 * Restates Shaders/StandardSurfaceBSDF.slang, Shaders/BSDFCommon.slang and
 * Shaders/Sampling.slang in FP32 C. */
#include "ora.h"
#include "ora_bsdf.h"

#include <string.h>

/* Material defaults, Source/MaterialBase.cpp:109-167 (Standard Surface defaults). */
void ora_material_defaults(ora_material_constants* m)
{
    memset(m, 0, sizeof(*m));
    m->base = 0.8f;
    m->baseColor[0] = m->baseColor[1] = m->baseColor[2] = 1.0f;
    m->diffuseRoughness = 0.0f;
    m->metalness = 0.0f;
    m->specular = 1.0f;
    m->specularColor[0] = m->specularColor[1] = m->specularColor[2] = 1.0f;
    m->specularRoughness = 0.2f;
    m->specularIOR = 1.5f;
    m->specularAnisotropy = 0.0f;
    m->specularRotation = 0.0f;
    m->transmission = 0.0f;
    m->transmissionColor[0] = m->transmissionColor[1] = m->transmissionColor[2] = 1.0f;
    m->subsurface = 0.0f;
    m->subsurfaceColor[0] = m->subsurfaceColor[1] = m->subsurfaceColor[2] = 1.0f;
    m->subsurfaceRadius[0] = m->subsurfaceRadius[1] = m->subsurfaceRadius[2] = 1.0f;
    m->subsurfaceScale = 1.0f;
    m->subsurfaceAnisotropy = 0.0f;
    m->sheen = 0.0f;
    m->sheenColor[0] = m->sheenColor[1] = m->sheenColor[2] = 1.0f;
    m->sheenRoughness = 0.3f;
    m->coat = 0.0f;
    m->coatColor[0] = m->coatColor[1] = m->coatColor[2] = 1.0f;
    m->coatRoughness = 0.1f;
    m->coatAnisotropy = 0.0f;
    m->coatRotation = 0.0f;
    m->coatIOR = 1.5f;
    m->coatAffectColor = 0.0f;
    m->coatAffectRoughness = 0.0f;
    m->emission = 0.0f;
    m->emissionColor[0] = m->emissionColor[1] = m->emissionColor[2] = 1.0f;
    m->opacity[0] = m->opacity[1] = m->opacity[2] = 1.0f;
    m->thinWalled = 0;
}

/* evaluateDefaultMaterial without textures, DefaultMaterial.slang:130-239:
 * plain copy of the constants, metalColor = baseColor (:236). */
void ora_material_from_constants(const ora_material_constants* c, ora_material* m)
{
    m->base = c->base;
    m->baseColor = v3load(c->baseColor);
    m->diffuseRoughness = c->diffuseRoughness;
    m->metalness = c->metalness;
    m->specular = c->specular;
    m->specularColor = v3load(c->specularColor);
    m->specularRoughness = c->specularRoughness;
    m->specularIOR = c->specularIOR;
    m->specularAnisotropy = c->specularAnisotropy;
    m->specularRotation = c->specularRotation;
    m->transmission = c->transmission;
    m->transmissionColor = v3load(c->transmissionColor);
    m->subsurface = c->subsurface;
    m->subsurfaceColor = v3load(c->subsurfaceColor);
    m->sheen = c->sheen;
    m->sheenColor = v3load(c->sheenColor);
    m->sheenRoughness = c->sheenRoughness;
    m->coat = c->coat;
    m->coatColor = v3load(c->coatColor);
    m->coatRoughness = c->coatRoughness;
    m->coatAnisotropy = c->coatAnisotropy;
    m->coatRotation = c->coatRotation;
    m->coatIOR = c->coatIOR;
    m->coatAffectColor = c->coatAffectColor;
    m->coatAffectRoughness = c->coatAffectRoughness;
    m->emission = c->emission;
    m->emissionColor = v3load(c->emissionColor);
    m->opacity = v3load(c->opacity);
    m->thinWalled = c->thinWalled != 0;
    m->metalColor = m->baseColor;
}

/* ---------------- Sampling.slang ---------------- */

/* sampleUniformDirection, Sampling.slang:20-34 */
static v3 sample_uniform_direction(const float r[2], float* pdf)
{
    float cosTheta = 1.0f - 2.0f * r[1];
    float sinTheta = sqrtf(1.0f - cosTheta * cosTheta);
    float phi = 2.0f * ORA_PI * r[0];
    *pdf = 0.25f * ORA_PI_INV;
    return V3(sinTheta * cosf(phi), sinTheta * sinf(phi), cosTheta);
}

/* sampleHemisphere(random, normal, pdf), Sampling.slang:39-52 */
static v3 sample_hemisphere_n(const float r[2], v3 n, float* pdf)
{
    v3 d = sample_uniform_direction(r, pdf);
    d = normalize(add(n, d));
    *pdf = dot(n, d) * ORA_PI_INV;
    return d;
}

/* ---------------- BSDFCommon.slang ---------------- */

static float remap_pass(float r, float t) { return r / t; }          /* :77-80 */
static float remap_fail(float r, float t) { return (r - t) / (1.0f - t); } /* :83-86 */

static v3 world_to_tangent(v3 i, v3 X, v3 Y, v3 N) { return V3(dot(i, X), dot(i, Y), dot(i, N)); }
static v3 tangent_to_world(v3 i, v3 X, v3 Y, v3 N)
{
    return add(add(scale(X, i.x), scale(Y, i.y)), scale(N, i.z));
}

/* computeRefraction, BSDFCommon.slang:122-152 */
static v3 compute_refraction(v3 V, float ior, int* reflected)
{
    *reflected = 0;
    int entering = V.z > 0.0f;
    v3 N = V3(0.0f, 0.0f, entering ? 1.0f : -1.0f);
    float eta = entering ? 1.0f / ior : ior / 1.0f;
    v3 L = refract3(neg(V), N, eta);
    if (length3(L) == 0.0f) {
        L = V3(-V.x, -V.y, V.z);
        *reflected = 1;
    }
    return L;
}

/* evaluateGGXNDF, BSDFCommon.slang:245-250 */
static float ggx_ndf(v3 H, float ax, float ay)
{
    float ellipse = sqrf(H.x) / sqrf(ax) + sqrf(H.y) / sqrf(ay) + sqrf(H.z);
    return 1.0f / (ORA_PI * ax * ay * sqrf(ellipse));
}

/* lambda, BSDFCommon.slang:208-213 */
static float smith_lambda(v3 v, float ax, float ay)
{
    float squared = (sqrf(ax) * sqrf(v.x) + sqrf(ay) * sqrf(v.y)) / sqrf(v.z);
    return (-1.0f + sqrtf(1.0f + squared)) * 0.5f;
}

/* evaluateSmithG1, BSDFCommon.slang:217-226 */
static float smith_g1(v3 V, v3 H, float ax, float ay)
{
    if (dot(V, H) <= 0.0f) return 0.0f;
    return 1.0f / (1.0f + smith_lambda(V, ax, ay));
}

/* evaluateGGXSmithVNDF, BSDFCommon.slang:255-269 */
static float ggx_smith_vndf(v3 V, v3 H, float ax, float ay)
{
    float VdotH = dot(V, H);
    if (VdotH <= 0.0f) return 0.0f;
    float D = ggx_ndf(H, ax, ay);
    float G1 = smith_g1(V, H, ax, ay);
    return D * G1 * VdotH / fabsf(V.z);
}

/* sampleGGXSmithVNDF, BSDFCommon.slang:274-299 (Heitz 2018, appendix A) */
static v3 sample_ggx_vndf(v3 V, float ax, float ay, const float r[2])
{
    v3 Vh = normalize(V3(ax * V.x, ay * V.y, V.z));
    float lensq = Vh.x * Vh.x + Vh.y * Vh.y;
    v3 T1 = lensq > 0.0f ? divs(V3(-Vh.y, Vh.x, 0.0f), sqrtf(lensq)) : V3(1.0f, 0.0f, 0.0f);
    v3 T2 = cross(Vh, T1);
    float rr = sqrtf(r[0]);
    float phi = 2.0f * ORA_PI * r[1];
    float t1 = rr * cosf(phi);
    float t2 = rr * sinf(phi);
    float s = 0.5f * (1.0f + Vh.z);
    t2 = (1.0f - s) * sqrtf(1.0f - t1 * t1) + s * t2;
    v3 Nh = add(add(scale(T1, t1), scale(T2, t2)), scale(Vh, sqrtf(maxf(0.0f, 1.0f - t1 * t1 - t2 * t2))));
    return normalize(V3(ax * Nh.x, ay * Nh.y, maxf(0.0f, Nh.z)));
}

/* sampleGGXSmithBRDF, BSDFCommon.slang:331-367, with anisotropy 0 (the ground plane's
 * matte reflection): alpha = max(1e-4, roughness^2) (roughnessToAlpha, :201-210); the
 * visible normal reflects V into L; pdf = VNDF / (4 V.H); the value is
 * D G2 F / (4 |V.z| |L.z|) at the half vector of V and L (evaluateGGXSmithBRDF, :303-323),
 * black when L leaves V's hemisphere. */
v3 ora_ggx_smith_brdf_sample(v3 V, v3 f0, float roughness, const float r[2], v3* L, float* pdf)
{
    float a = maxf(0.0001f, sqrf(roughness));
    int flip = V.z <= 0.0f;
    v3 H = sample_ggx_vndf(flip ? neg(V) : V, a, a, r);
    if (flip) H = neg(H);
    *L = reflect3(neg(V), H);
    *pdf = ggx_smith_vndf(V, H, a, a) / (4.0f * dot(V, H));
    if (!(V.z * L->z > 0.0f)) return v3s(0.0f);
    v3 hr = normalize(add(V, *L));
    float D = ggx_ndf(hr, a, a);
    float G2 = (dot(V, hr) <= 0.0f || dot(*L, hr) < 0.0f)
                   ? 0.0f
                   : 1.0f / (1.0f + smith_lambda(V, a, a) + smith_lambda(*L, a, a));
    v3 F = lerp3(f0, v3s(1.0f), powf(1.0f - dot(V, hr), 5.0f));
    return divs(scale(F, D * G2), 4.0f * fabsf(V.z) * fabsf(L->z));
}

/* ---------------- StandardSurfaceBSDF.slang ---------------- */

static int same_hemisphere_n(v3 a, v3 b, v3 n) { return dot(a, n) * dot(b, n) > 0.0f; } /* :22-25 */

/* mx_roughness_anisotropy, :31-56 */
static void mx_roughness_anisotropy(float roughness, float anisotropy, float out[2])
{
    float rs = clampf(roughness * roughness, ORA_FLOAT_EPS, 1.0f);
    if (anisotropy > 0.0f) {
        float aspect = sqrtf(1.0f - clampf(anisotropy, 0.0f, 0.98f));
        out[0] = minf(rs / aspect, 1.0f);
        out[1] = rs * aspect;
    } else {
        out[0] = rs;
        out[1] = rs;
    }
}

static float mx_pow5(float x) { return sqrf(sqrf(x)) * x; }

/* mx_mix, :71-74 */
static v3 mx_mix(v3 bg, v3 fg, float w) { return lerp3(bg, fg, clampf(w, 0.0f, 1.0f)); }

/* mx_orennayar, :76-89 */
static float mx_orennayar(v3 L, v3 V, v3 N, float NdotL, float roughness)
{
    float LdotV = dot(L, V);
    float NdotV = fabsf(dot(N, V));
    float t = LdotV - NdotL * NdotV;
    t = t > 0.0f ? t / maxf(NdotL, NdotV) : 0.0f;
    float sigma2 = sqrf(roughness * ORA_PI);
    float A = 1.0f - 0.5f * (sigma2 / (sigma2 + 0.33f));
    float B = 0.45f * sigma2 / (sigma2 + 0.09f);
    return A + B * t;
}

/* mx_diffuse_brdf_reflection, :91-112 */
static v3 mx_diffuse_brdf_reflection(v3 L, v3 V, float weight, v3 color, float roughness, v3 N)
{
    if (weight < ORA_FLOAT_EPS || !same_hemisphere_n(L, V, N)) return v3s(0.0f);
    float NdotL = maxf(dot(L, N), 0.0f);
    v3 result = scale(color, weight * NdotL * ORA_PI_INV);
    if (roughness > 0.0f) result = scale(result, mx_orennayar(L, V, N, NdotL, roughness));
    return result;
}

static float mx_average_roughness(const float r[2]) { return sqrtf(r[0] * r[1]); } /* :115-118 */
static float mx_ior_to_f0(float ior) { return sqrf((ior - 1.0f) / (ior + 1.0f)); }   /* :121-124 */

/* mx_ggx_NDF, :128-134 */
static float mx_ggx_ndf(v3 X, v3 Y, v3 H, float NdotH, float ax, float ay)
{
    float XdotH = dot(X, H);
    float YdotH = dot(Y, H);
    float denom = sqrf(XdotH / ax) + sqrf(YdotH / ay) + sqrf(NdotH);
    return 1.0f / (ORA_PI * ax * ay * sqrf(denom));
}

/* mx_fresnel_schlick, :137-142 */
static float mx_fresnel_schlick(float cosTheta, float F0)
{
    float x = clampf(1.0f - cosTheta, 0.0f, 1.0f);
    float x5 = mx_pow5(x);
    return F0 + (1.0f - F0) * x5;
}

/* mx_ggx_smith_G, :146-152 */
static float mx_ggx_smith_g(float NdotL, float NdotV, float alpha)
{
    float alpha2 = sqrf(alpha);
    float lambdaL = sqrtf(alpha2 + (1.0f - alpha2) * sqrf(NdotL));
    float lambdaV = sqrtf(alpha2 + (1.0f - alpha2) * sqrf(NdotV));
    return 2.0f / (lambdaL / NdotL + lambdaV / NdotV);
}

/* mx_ggx_directional_albedo_curve_fit, :158-166 (scalar F0/F90 case, .x component) */
static float mx_ggx_dir_albedo(float NdotV, float roughness, float F0, float F90)
{
    float r0 = roughness * -1.0f + 1.0f;
    float r1 = roughness * -0.0275f + 0.0425f;
    float r2 = roughness * -0.572f + 1.04f;
    float r3 = roughness * 0.022f + -0.04f;
    float a004 = minf(r0 * r0, exp2f(-9.28f * NdotV)) * r0 + r1;
    float ABx = -1.04f * a004 + r2;
    float ABy = 1.04f * a004 + r3;
    return F0 * ABx + F90 * ABy;
}

/* mx_ggx_energy_compensation, :175-184 */
static float mx_ggx_energy_comp(float NdotV, float roughness, float Fss)
{
    float Ess = mx_ggx_dir_albedo(NdotV, roughness, 1.0f, 1.0f);
    return 1.0f + Fss * (1.0f - Ess) / Ess;
}
static v3 mx_ggx_energy_comp3(float NdotV, float roughness, v3 Fss)
{
    float Ess = mx_ggx_dir_albedo(NdotV, roughness, 1.0f, 1.0f);
    return add(v3s(1.0f), divs(scale(Fss, 1.0f - Ess), Ess));
}

static v3 fabs3(v3 a) { return V3(fabsf(a.x), fabsf(a.y), fabsf(a.z)); }

/* mx_dielectric_brdf_reflection, :186-220 */
static v3 mx_dielectric_brdf_reflection(v3 L, v3 V, float weight, v3 tint, float ior,
                                        const float rough[2], v3 N, v3 X, v3 base)
{
    if (weight < ORA_FLOAT_EPS) return base;
    float NdotL = fabsf(dot(N, L));
    float NdotV = fabsf(dot(N, V));
    v3 Y = normalize(cross(N, X));
    v3 H = normalize(add(L, V));
    float NdotH = dot(N, H);
    float VdotH = dot(V, H);
    float avgR = mx_average_roughness(rough);
    float F0 = mx_ior_to_f0(ior);
    float D = mx_ggx_ndf(X, Y, H, NdotH, rough[0], rough[1]);
    float F = mx_fresnel_schlick(VdotH, F0);
    float G = mx_ggx_smith_g(NdotL, NdotV, avgR);
    float comp = mx_ggx_energy_comp(NdotV, avgR, F);
    float dirAlbedo = mx_ggx_dir_albedo(NdotV, avgR, F0, 1.0f) * comp;
    v3 result = fabs3(divs(scale(tint, D * F * G * comp * weight), 4.0f * NdotV));
    return add(result, scale(base, 1.0f - dirAlbedo * weight));
}

/* mx_artistic_to_complex_ior, :222-236 */
static void mx_artistic_to_complex_ior(v3 refl, v3 edge, v3* n, v3* k)
{
    v3 r = V3(clampf(refl.x, 0.0f, 0.99f), clampf(refl.y, 0.0f, 0.99f), clampf(refl.z, 0.0f, 0.99f));
    v3 rs = V3(sqrtf(r.x), sqrtf(r.y), sqrtf(r.z));
    v3 nmin = divv(sub(v3s(1.0f), r), add(v3s(1.0f), r));
    v3 nmax = divv(add(v3s(1.0f), rs), sub(v3s(1.0f), rs));
    v3 ior = lerp3v(nmax, nmin, edge);
    v3 np1 = add(ior, v3s(1.0f));
    v3 nm1 = sub(ior, v3s(1.0f));
    v3 k2 = divv(sub(mul(mul(np1, np1), r), mul(nm1, nm1)), sub(v3s(1.0f), r));
    k2 = V3(maxf(k2.x, 0.0f), maxf(k2.y, 0.0f), maxf(k2.z, 0.0f));
    *n = ior;
    *k = V3(sqrtf(k2.x), sqrtf(k2.y), sqrtf(k2.z));
}

/* mx_fresnel_conductor, :238-250 */
static v3 mx_fresnel_conductor(float cosTheta, v3 n, v3 k)
{
    float c2 = cosTheta * cosTheta;
    v3 n2k2 = add(mul(n, n), mul(k, k));
    v3 nc2 = scale(n, 2.0f * cosTheta);
    v3 rs_a = add(n2k2, v3s(c2));
    v3 rp_a = add(scale(n2k2, c2), v3s(1.0f));
    v3 rs = divv(sub(rs_a, nc2), add(rs_a, nc2));
    v3 rp = divv(sub(rp_a, nc2), add(rp_a, nc2));
    return scale(add(rs, rp), 0.5f);
}

/* mx_conductor_brdf_reflection, :252-285 */
static v3 mx_conductor_brdf_reflection(v3 L, v3 V, float weight, v3 refl, v3 edge,
                                       const float rough[2], v3 N, v3 X)
{
    if (weight < ORA_FLOAT_EPS || !same_hemisphere_n(L, V, N)) return v3s(0.0f);
    float NdotL = fabsf(dot(N, L));
    float NdotV = fabsf(dot(N, V));
    v3 Y = normalize(cross(N, X));
    v3 H = normalize(add(L, V));
    float NdotH = dot(N, H);
    float VdotH = dot(V, H);
    v3 n, k;
    mx_artistic_to_complex_ior(refl, edge, &n, &k);
    float avgR = mx_average_roughness(rough);
    float D = mx_ggx_ndf(X, Y, H, NdotH, rough[0], rough[1]);
    v3 F = mx_fresnel_conductor(VdotH, n, k);
    float G = mx_ggx_smith_g(NdotL, NdotV, avgR);
    v3 comp = mx_ggx_energy_comp3(NdotV, avgR, F);
    return divs(scale(mul(F, comp), D * G * weight), 4.0f * NdotV);
}

/* mx_imageworks_sheen_NDF, :292-299 */
static float mx_sheen_ndf(float cosTheta, float roughness)
{
    float inv = 1.0f / roughness;
    float cos2 = cosTheta * cosTheta;
    float sin2 = 1.0f - cos2;
    return (2.0f + inv) * powf(sin2, inv * 0.5f) / (2.0f * ORA_PI);
}

/* Sheen directional-albedo table, :307-337 (data from MaterialX mx_sheen_brdf.glsl). */
static const float kSheenAlbedo[256] = { 1.6177f,
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
    0.147049f };

/* mx_imageworks_sheen_directional_albedo, :341-362 */
float ora_sheen_albedo(float cosTheta, float roughness)
{
    float x = cosTheta * 15.0f;
    float y = roughness * 15.0f;
    /* NaN reads row/column 0 and indices stay inside the table (the GPU path does
     * the same); finite arguments in [0, 1] are unaffected. */
    if (!(x >= 0.0f)) x = 0.0f;
    if (!(y >= 0.0f)) y = 0.0f;
    int ix = (int)fminf(x, 15.0f), iy = (int)fminf(y, 15.0f);
    int ix2 = ix + 1 < 0 ? 0 : (ix + 1 > 15 ? 15 : ix + 1);
    int iy2 = iy + 1 < 0 ? 0 : (iy + 1 > 15 ? 15 : iy + 1);
    float fx = x - (float)ix, fy = y - (float)iy;
    float v1 = lerpf(kSheenAlbedo[iy * 16 + ix], kSheenAlbedo[iy * 16 + ix2], fx);
    float v2 = lerpf(kSheenAlbedo[iy2 * 16 + ix], kSheenAlbedo[iy2 * 16 + ix2], fx);
    return clampf(lerpf(v1, v2, fy), 0.0f, 1.0f);
}

/* mx_sheen_brdf_reflection, :366-410 */
static v3 mx_sheen_brdf_reflection(v3 L, v3 V, float weight, v3 color, float roughness, v3 N, v3 base)
{
    if (weight < ORA_FLOAT_EPS) return base;
    if (dot(N, V) < 0.0f) N = neg(N);
    v3 H = normalize(add(L, V));
    float NdotL = clampf(dot(N, L), ORA_FLOAT_EPS, 1.0f);
    float NdotV = clampf(dot(N, V), ORA_FLOAT_EPS, 1.0f);
    float NdotH = clampf(dot(N, H), ORA_FLOAT_EPS, 1.0f);
    float alpha = clampf(roughness, ORA_FLOAT_EPS, 1.0f);
    float D = mx_sheen_ndf(NdotH, alpha);
    v3 fr = divs(scale(color, D), 4.0f * (NdotL + NdotV - NdotL * NdotV));
    float dirAlbedo = ora_sheen_albedo(NdotV, alpha);
    return add(scale(fr, NdotL * weight), scale(base, 1.0f - dirAlbedo * weight));
}

/* mx_sheen_brdf_specular_layer, :414-458 */
static v3 mx_sheen_brdf_specular_layer(v3 L, v3 V, float weight, v3 color, float roughness, v3 N)
{
    if (!same_hemisphere_n(L, V, N)) return v3s(0.0f);
    if (dot(N, V) < 0.0f) N = neg(N);
    v3 H = normalize(add(L, V));
    float NdotL = clampf(dot(N, L), ORA_FLOAT_EPS, 1.0f);
    float NdotV = clampf(dot(N, V), ORA_FLOAT_EPS, 1.0f);
    float NdotH = clampf(dot(N, H), ORA_FLOAT_EPS, 1.0f);
    float alpha = clampf(roughness, ORA_FLOAT_EPS, 1.0f);
    float D = mx_sheen_ndf(NdotH, alpha);
    v3 fr = divs(scale(color, D), 4.0f * (NdotL + NdotV - NdotL * NdotV));
    return scale(fr, NdotL * weight);
}

/* mx_subsurface_brdf_reflection, :462-473 */
static v3 mx_subsurface_brdf_reflection(v3 L, v3 V, float weight, v3 color, v3 N)
{
    (void)V;
    float NdotL = dot(L, neg(N));
    if (NdotL <= 0.0f || weight < ORA_FLOAT_EPS) return v3s(0.0f);
    return scale(color, weight * NdotL * ORA_PI_INV);
}

/* mx_rotate_vector3 with mx_rotationMatrix, :477-502.  The float4x4 constructor
 * fills rows; mul(m, v) is row . v. */
static v3 mx_rotate_vector3(v3 vec, float amount, v3 axis)
{
    float angle = amount * (ORA_PI / 180.0f);
    axis = normalize(axis);
    float s = sinf(angle), c = cosf(angle), oc = 1.0f - c;
    float m00 = oc * axis.x * axis.x + c, m01 = oc * axis.x * axis.y - axis.z * s, m02 = oc * axis.z * axis.x + axis.y * s;
    float m10 = oc * axis.x * axis.y + axis.z * s, m11 = oc * axis.y * axis.y + c, m12 = oc * axis.y * axis.z - axis.x * s;
    float m20 = oc * axis.z * axis.x - axis.y * s, m21 = oc * axis.y * axis.z + axis.x * s, m22 = oc * axis.z * axis.z + c;
    return V3(m00 * vec.x + m01 * vec.y + m02 * vec.z,
              m10 * vec.x + m11 * vec.y + m12 * vec.z,
              m20 * vec.x + m21 * vec.y + m22 * vec.z);
}

/* mx_dielectric_brdf_diffuse_layer, :506-540 */
static v3 mx_dielectric_brdf_diffuse_layer(v3 L, v3 V, float weight, float ior, const float rough[2],
                                           v3 N, v3 albedo)
{
    if (weight < ORA_FLOAT_EPS || !same_hemisphere_n(L, V, N)) return v3s(0.0f);
    float NdotV = fabsf(dot(N, V));
    v3 H = normalize(add(L, V));
    float VdotH = dot(V, H);
    float avgR = mx_average_roughness(rough);
    float F0 = mx_ior_to_f0(ior);
    float F = mx_fresnel_schlick(VdotH, F0);
    float comp = mx_ggx_energy_comp(NdotV, avgR, F);
    float dirAlbedo = mx_ggx_dir_albedo(NdotV, avgR, F0, 1.0f) * comp;
    v3 result = v3s(1.0f - dirAlbedo);
    return scale(mul(result, albedo), weight * ORA_PI_INV);
}

/* mx_dielectric_brdf_specular_layer, :545-584 */
static v3 mx_dielectric_brdf_specular_layer(v3 L, v3 V, v3 tint, float ior, const float rough[2], v3 N, v3 X)
{
    if (!same_hemisphere_n(L, V, N)) return v3s(0.0f);
    float NdotL = fabsf(dot(N, L));
    float NdotV = fabsf(dot(N, V));
    v3 Y = normalize(cross(N, X));
    v3 H = normalize(add(L, V));
    float NdotH = dot(N, H);
    float VdotH = dot(V, H);
    float avgR = mx_average_roughness(rough);
    float F0 = mx_ior_to_f0(ior);
    float D = mx_ggx_ndf(X, Y, H, NdotH, rough[0], rough[1]);
    float F = mx_fresnel_schlick(VdotH, F0);
    float G = mx_ggx_smith_g(NdotL, NdotV, avgR);
    float comp = mx_ggx_energy_comp(NdotV, avgR, F);
    return fabs3(divs(scale(tint, D * F * G * comp), 4.0f * NdotV * NdotL));
}

/* Values shared by evaluate/sample: the common prologue of :598-629, :862-905, :1093-1143. */
typedef struct {
    float main_roughness[2];
    v3 main_tangent, coat_tangent;
    v3 coat_affected_diffuse_color, coat_affected_subsurface_color;
} prologue;

static void make_prologue(const ora_material* m, const ora_shading* sh, prologue* p)
{
    float caf = m->coatAffectRoughness * m->coat * m->coatRoughness;
    float car = lerpf(m->specularRoughness, 1.0f, caf);
    float coat_clamped = clampf(m->coat, 0.0f, 1.0f);
    float coat_gamma = coat_clamped * m->coatAffectColor + 1.0f;
    p->coat_affected_subsurface_color = pow3(m->subsurfaceColor, coat_gamma);
    mx_roughness_anisotropy(car, m->specularAnisotropy, p->main_roughness);
    p->main_tangent = mx_rotate_vector3(sh->tangent, m->specularRotation * 360.0f, sh->normal);
    p->coat_tangent = mx_rotate_vector3(sh->tangent, m->coatRotation * 360.0f, sh->normal);
    p->coat_affected_diffuse_color = pow3(m->baseColor, coat_gamma);
}

/* evaluateMaterial, :598-718 */
v3 ora_eval_material(const ora_material* m, const ora_shading* sh, v3 Vw, v3 Lw)
{
    prologue p;
    make_prologue(m, sh, &p);
    v3 N = sh->normal;
    v3 diffuse = mx_diffuse_brdf_reflection(Lw, Vw, m->base, p.coat_affected_diffuse_color,
                                            m->diffuseRoughness, N);
    v3 subsurface_mix = diffuse;
    if (m->subsurface > 0.0f) {
        v3 ss = mx_subsurface_brdf_reflection(Lw, Vw, 1.0f, p.coat_affected_subsurface_color, N);
        subsurface_mix = mx_mix(diffuse, ss, m->subsurface);
    }
    if (dot(N, Vw) * dot(N, Lw) < 0.0f) return subsurface_mix;
    v3 sheen_and_diffuse = subsurface_mix;
    if (m->sheen > 0.0f)
        sheen_and_diffuse = mx_sheen_brdf_reflection(Lw, Vw, m->sheen, m->sheenColor, m->sheenRoughness,
                                                     N, subsurface_mix);
    sheen_and_diffuse = scale(sheen_and_diffuse, 1.0f - m->transmission);
    v3 specular = sheen_and_diffuse;
    if (m->specular > 0.0f)
        specular = mx_dielectric_brdf_reflection(Lw, Vw, m->specular, m->specularColor, m->specularIOR,
                                                 p.main_roughness, N, p.main_tangent, sheen_and_diffuse);
    v3 metalness_mix = specular;
    if (m->metalness > 0.0f) {
        v3 refl = scale(m->metalColor, m->base);
        v3 edge = scale(m->specularColor, m->specular);
        v3 metal = mx_conductor_brdf_reflection(Lw, Vw, 1.0f, refl, edge, p.main_roughness, N, p.main_tangent);
        metalness_mix = mx_mix(specular, metal, m->metalness);
    }
    v3 final_color = metalness_mix;
    if (m->coat > 0.0f) {
        float coat_rough[2];
        mx_roughness_anisotropy(m->coatRoughness, m->coatAnisotropy, coat_rough);
        v3 att = lerp3(v3s(1.0f), m->coatColor, m->coat);
        v3 mma = mul(att, metalness_mix);
        final_color = mx_dielectric_brdf_reflection(Lw, Vw, m->coat, m->coatColor, m->coatIOR, coat_rough,
                                                    N, p.coat_tangent, mma);
    }
    return final_color;
}

/* calculateGGXPDF, :737-746 */
static float calc_ggx_pdf(v3 V, const float alpha[2], v3 L)
{
    v3 H = normalize(add(V, L));
    float denom = 4.0f * dot(V, H);
    return ggx_smith_vndf(V, H, alpha[0], alpha[1]) / denom;
}

/* calculateGGXLightDirectionAndPDF, :751-773 */
static v3 calc_ggx_dir_pdf(v3 X, v3 Y, v3 N, v3 V, const float alpha[2], const float r[2], v3* Lw, float* pdf)
{
    int flip = V.z <= 0.0f;
    v3 H = sample_ggx_vndf(flip ? neg(V) : V, alpha[0], alpha[1], r);
    if (flip) H = neg(H);
    v3 L = reflect3(neg(V), H);
    float denom = 4.0f * dot(V, H);
    *pdf = ggx_smith_vndf(V, H, alpha[0], alpha[1]) / denom;
    *Lw = tangent_to_world(L, X, Y, N);
    return L;
}

/* evaluateMaterialAndPDF, :862-1088.  Progressive path (no denoiser AOVs):
 * getGlossySelectionProbability returns specLuminance and the non-glossy
 * compensation is exactly 1 (:792-834). */
v3 ora_eval_material_pdf(const ora_material* m, const ora_shading* sh, v3 Vw, v3 Lw, const float rin[2],
                         float* pdf, int* lobe)
{
    v3 N = sh->normal, X = sh->tangent, Y = sh->bitangent;
    v3 V = world_to_tangent(Vw, X, Y, N);
    v3 L = world_to_tangent(Lw, X, Y, N);
    float NdotV = fabsf(V.z);
    float r[2] = { rin[0], rin[1] };
    *pdf = 0.0f;
    *lobe = ORA_NO_SECONDARY_RAY;
    prologue p;
    make_prologue(m, sh, &p);
    v3 result = v3s(0.0f);
    float coat = m->coat;
    float coatF0 = mx_ior_to_f0(m->coatIOR);
    float coatF = mx_fresnel_schlick(NdotV, coatF0);
    float coatLum = luminance(scale(m->coatColor, coatF * coat));
    if (r[0] < coatLum) {
        r[0] = remap_pass(r[0], coatLum);
        float ca[2];
        mx_roughness_anisotropy(m->coatRoughness, m->coatAnisotropy, ca);
        *pdf = calc_ggx_pdf(V, ca, L);
        result = mx_dielectric_brdf_specular_layer(Lw, Vw, scale(m->coatColor, m->coat), m->coatIOR, ca, N,
                                                   p.coat_tangent);
        *lobe = ORA_COAT_LOBE;
    } else {
        r[0] = remap_fail(r[0], coatLum);
        float sa[2];
        mx_roughness_anisotropy(m->specularRoughness, m->specularAnisotropy, sa);
        if (r[0] < m->metalness) {
            r[0] = remap_pass(r[0], m->metalness);
            *pdf = calc_ggx_pdf(V, sa, L);
            v3 refl = scale(m->metalColor, m->base);
            v3 edge = scale(m->specularColor, m->specular);
            result = mx_conductor_brdf_reflection(Lw, Vw, 1.0f, refl, edge, p.main_roughness, N, p.main_tangent);
            result = divs(result, fabsf(L.z));
            *lobe = ORA_METAL_LOBE;
        } else {
            r[0] = remap_fail(r[0], m->metalness);
            float F0 = mx_ior_to_f0(m->specularIOR);
            float F = mx_fresnel_schlick(NdotV, F0);
            float specLum = luminance(scale(m->specularColor, F * m->specular));
            float gsp = specLum;
            if (r[0] < gsp) {
                r[0] = remap_pass(r[0], gsp);
                *pdf = calc_ggx_pdf(V, sa, L);
                result = mx_dielectric_brdf_specular_layer(Lw, Vw, scale(m->specularColor, m->specular),
                                                           m->specularIOR, p.main_roughness, sh->normal,
                                                           p.main_tangent);
                result = divs(result, gsp);
                *lobe = ORA_GLOSSY_LOBE;
            } else {
                r[0] = remap_fail(r[0], gsp);
                float sheenAlpha = clampf(m->sheenRoughness, ORA_FLOAT_EPS, 1.0f);
                float sheen = m->sheen;
                if (r[0] < m->sheen) {
                    r[0] = remap_pass(r[0], sheen);
                    *pdf = dot(N, Lw) * ORA_PI_INV;
                    v3 H = normalize(add(L, V));
                    result = mx_sheen_brdf_specular_layer(Lw, Vw, m->sheen, m->sheenColor, m->sheenRoughness, N);
                    *pdf = mx_sheen_ndf(H.z, sheenAlpha);
                    *lobe = ORA_SHEEN_LOBE;
                } else {
                    r[0] = remap_fail(r[0], sheen);
                    *pdf = dot(N, Lw) * ORA_PI_INV;
                    result = mx_dielectric_brdf_diffuse_layer(Lw, Vw, m->base, m->specularIOR, p.main_roughness,
                                                              N, p.coat_affected_diffuse_color);
                    float dirAlbedo = ora_sheen_albedo(NdotV, sheenAlpha);
                    result = scale(result, 1.0f - dirAlbedo * m->sheen);
                    *lobe = ORA_DIFFUSE_LOBE;
                }
            }
        }
        v3 att = lerp3(v3s(1.0f), m->coatColor, m->coat);
        result = mul(result, att);
    }
    float NdotL = fabsf(L.z);
    return scale(result, NdotL);
}

/* sampleMaterial, :1093-1372 */
v3 ora_sample_material(const ora_material* m, const ora_shading* sh, v3 Vw, const float rin[2],
                       v3* Lw_out, float* pdf, int* lobe)
{
    v3 L = V3(0.0f, 0.0f, 0.0f);
    v3 N = sh->normal, X = sh->tangent, Y = sh->bitangent;
    v3 V = world_to_tangent(Vw, X, Y, N);
    float NdotV = fabsf(V.z);
    float r[2] = { rin[0], rin[1] };
    v3 Lw = V3(1.0f, 0.0f, 0.0f);
    *pdf = 0.0f;
    *lobe = ORA_NO_SECONDARY_RAY;
    prologue p;
    make_prologue(m, sh, &p);
    v3 result = v3s(0.0f);
    float coat = m->coat;
    float coatF0 = mx_ior_to_f0(m->coatIOR);
    float coatF = mx_fresnel_schlick(NdotV, coatF0);
    float coatLum = luminance(scale(m->coatColor, coatF * coat));
    if (r[0] < coatLum) {
        r[0] = remap_pass(r[0], coatLum);
        float ca[2];
        mx_roughness_anisotropy(m->coatRoughness, m->coatAnisotropy, ca);
        L = calc_ggx_dir_pdf(X, Y, N, V, ca, r, &Lw, pdf);
        result = mx_dielectric_brdf_specular_layer(Lw, Vw, scale(m->coatColor, coat), m->coatIOR, ca, N,
                                                   p.coat_tangent);
        result = divs(result, coatLum);
        *lobe = ORA_COAT_LOBE;
    } else {
        r[0] = remap_fail(r[0], coatLum);
        float sa[2];
        mx_roughness_anisotropy(m->specularRoughness, m->specularAnisotropy, sa);
        if (r[0] < m->metalness) {
            r[0] = remap_pass(r[0], m->metalness);
            L = calc_ggx_dir_pdf(X, Y, N, V, sa, r, &Lw, pdf);
            v3 refl = scale(m->metalColor, m->base);
            v3 edge = scale(m->specularColor, m->specular);
            result = mx_conductor_brdf_reflection(Lw, Vw, 1.0f, refl, edge, p.main_roughness, N, p.main_tangent);
            result = divs(result, fabsf(L.z));
            *lobe = ORA_METAL_LOBE;
        } else {
            r[0] = remap_fail(r[0], m->metalness);
            float F0 = mx_ior_to_f0(m->specularIOR);
            float F = mx_fresnel_schlick(NdotV, F0);
            float specLum = luminance(scale(m->specularColor, F * m->specular));
            float gsp = specLum;
            if (r[0] < gsp) {
                r[0] = remap_pass(r[0], gsp);
                L = calc_ggx_dir_pdf(X, Y, N, V, sa, r, &Lw, pdf);
                result = mx_dielectric_brdf_specular_layer(Lw, Vw, scale(m->specularColor, m->specular),
                                                           m->specularIOR, p.main_roughness, sh->normal,
                                                           p.main_tangent);
                result = divs(result, gsp);
                *lobe = ORA_GLOSSY_LOBE;
            } else {
                r[0] = remap_fail(r[0], gsp);
                float sheenAlpha = clampf(m->sheenRoughness, ORA_FLOAT_EPS, 1.0f);
                /* :1268 gates on ENABLE_COAT (both flags are 1) */
                float sheen = m->sheen;
                if (r[0] < m->sheen) {
                    r[0] = remap_pass(r[0], sheen);
                    v3 Nh = V3(0.0f, 0.0f, V.z > 0.0f ? 1.0f : -1.0f);
                    L = sample_hemisphere_n(r, Nh, pdf);
                    Lw = tangent_to_world(L, X, Y, N);
                    v3 H = normalize(add(L, V));
                    result = mx_sheen_brdf_specular_layer(Lw, Vw, m->sheen, m->sheenColor, m->sheenRoughness, N);
                    *pdf = mx_sheen_ndf(H.z, sheenAlpha);
                    *lobe = ORA_SHEEN_LOBE;
                } else {
                    r[0] = remap_fail(r[0], sheen);
                    float transmission = m->transmission;
                    if (r[0] < transmission) {
                        r[0] = remap_pass(r[0], transmission);
                        int reflected = 0;
                        L = neg(V);
                        if (!m->thinWalled) L = compute_refraction(V, m->specularIOR, &reflected);
                        Lw = tangent_to_world(L, X, Y, N);
                        result = divs(m->transmissionColor, fabsf(L.z));
                        *pdf = 1.0f;
                        *lobe = reflected ? ORA_GLOSSY_LOBE : ORA_TRANSMISSION_LOBE;
                    } else {
                        r[0] = remap_fail(r[0], transmission);
                        v3 Nh = V3(0.0f, 0.0f, V.z > 0.0f ? 1.0f : -1.0f);
                        L = sample_hemisphere_n(r, Nh, pdf);
                        Lw = tangent_to_world(L, X, Y, N);
                        result = mx_dielectric_brdf_diffuse_layer(Lw, Vw, m->base, m->specularIOR,
                                                                  p.main_roughness, N,
                                                                  p.coat_affected_diffuse_color);
                        float dirAlbedo = ora_sheen_albedo(NdotV, sheenAlpha);
                        result = scale(result, 1.0f - dirAlbedo * m->sheen);
                        *lobe = ORA_DIFFUSE_LOBE;
                    }
                }
            }
        }
        v3 att = lerp3(v3s(1.0f), m->coatColor, m->coat);
        result = mul(result, att);
    }
    float NdotL = fabsf(L.z);
    *Lw_out = Lw;
    return scale(result, NdotL);
}

/* ---- exported wrappers on an explicit frame (for unit tests) ---- */
static void frame_to_shading(const ora_frame* f, ora_shading* sh)
{
    memset(sh, 0, sizeof(*sh));
    sh->normal = v3load(f->N);
    sh->tangent = v3load(f->X);
    sh->bitangent = v3load(f->Y);
}

void ora_bsdf_evaluate(const ora_material_constants* mc, const ora_frame* f, const float V[3],
                       const float L[3], float out[3])
{
    ora_material m;
    ora_shading sh;
    ora_material_from_constants(mc, &m);
    frame_to_shading(f, &sh);
    v3store(out, ora_eval_material(&m, &sh, v3load(V), v3load(L)));
}

void ora_bsdf_evaluate_pdf(const ora_material_constants* mc, const ora_frame* f, const float V[3],
                           const float L[3], const float rnd[2], float out[3], float* pdf, int* lobe)
{
    ora_material m;
    ora_shading sh;
    ora_material_from_constants(mc, &m);
    frame_to_shading(f, &sh);
    v3store(out, ora_eval_material_pdf(&m, &sh, v3load(V), v3load(L), rnd, pdf, lobe));
}

void ora_bsdf_sample(const ora_material_constants* mc, const ora_frame* f, const float V[3],
                     const float rnd[2], float out[3], float L[3], float* pdf, int* lobe)
{
    ora_material m;
    ora_shading sh;
    ora_material_from_constants(mc, &m);
    frame_to_shading(f, &sh);
    v3 Lw;
    v3store(out, ora_sample_material(&m, &sh, v3load(V), rnd, &Lw, pdf, lobe));
    v3store(L, Lw);
}
