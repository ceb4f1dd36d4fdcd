/* Oracle: the path tracer, one pixel-sample at a time (test infrastructure only).
 * Restates Shaders/MainEntryPoints.slang (RayGenShader, miss shaders),
 * Shaders/HitShaderEntryPoints.slang (closest hit), Shaders/ShadeFunctions.slang,
 * Shaders/RayTrace.slang (camera), Shaders/Geometry.slang (shading data) and
 * HGI/Shaders/Accumulation.glsl. */
#include "ora.h"
#include "ora_bsdf.h"
#include "ora_geom.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* computeShadingPosition, Geometry.slang:131-148 */
static v3 shading_position(v3 geom, v3 shadingNormal, const v3 p[3], const v3 n[3], const float b[3])
{
    v3 pp[3];
    for (int i = 0; i < 3; i++) pp[i] = sub(geom, scale(n[i], dot(sub(geom, p[i]), n[i])));
    v3 sp = add(add(scale(pp[0], b[0]), scale(pp[1], b[1])), scale(pp[2], b[2]));
    int convex = dot(sub(sp, geom), shadingNormal) > 0.0f;
    return convex ? sp : geom;
}

/* computeShadingData, Geometry.slang:152-265 (no tangents on HGI: HGIScene.h:92). */
static void shading_data(const ora_scene* s, uint32_t ii, uint32_t prim, float bu, float bv, ora_shading* sh)
{
    const ora_instance* in = &s->inst[ii];
    const ora_mesh* me = &s->meshes[in->mesh];
    float b[3] = { 1.0f - bu - bv, bu, bv };
    const uint32_t* tri = me->idx + 3 * (size_t)prim;
    v3 p[3], n[3];
    v3 geom = v3s(0.0f), nrm = v3s(0.0f);
    sh->texCoord[0] = sh->texCoord[1] = 0.0f;
    for (int i = 0; i < 3; i++) {
        p[i] = v3load(me->pos + 3 * (size_t)tri[i]);
        geom = add(geom, scale(p[i], b[i]));
        /* texCoord, Geometry.slang:199-203 */
        sh->texCoord[0] = sh->texCoord[0] + me->uv[2 * (size_t)tri[i]] * b[i];
        sh->texCoord[1] = sh->texCoord[1] + me->uv[2 * (size_t)tri[i] + 1] * b[i];
        if (me->nrm) {
            n[i] = v3load(me->nrm + 3 * (size_t)tri[i]);
            nrm = add(nrm, scale(n[i], b[i]));
        }
    }
    v3 objN = normalize(me->nrm ? nrm : cross(sub(p[1], p[0]), sub(p[2], p[0])));
    v3 objP = me->nrm ? shading_position(geom, nrm, p, n, b) : geom;
    sh->position = xform_point(in->o2w, objP);
    sh->geomPosition = xform_point(in->o2w, geom);
    sh->normal = normalize(xform_vector(in->o2w, objN));
    v3 up = fabsf(sh->normal.y) < 0.999f ? V3(0.0f, 1.0f, 0.0f) : V3(1.0f, 0.0f, 0.0f);
    sh->tangent = normalize(cross(sh->normal, up));
    sh->bitangent = cross(sh->normal, sh->tangent);
}

/* buildBasis, Globals.slang:205-210 */
static void build_basis(v3 n, v3* t, v3* b);

/* The textured part of evaluateDefaultMaterial (DefaultMaterial.slang:171-233) for
 * material mi: base colour (and the metal colour copied from it), specular
 * roughness and emission colour into m; with a normal map, the tangent-space
 * normal (calculateNormalFromMap, :107-127) replaces the shading normal and frame
 * (HitShaderEntryPoints.slang:251-255).  Opacity is applied by the layer loop. */
static void material_textures(const ora_scene* s, const ora_material_constants* mc, uint32_t mi, ora_material* m,
                              ora_shading* sh)
{
    float t[3];
    if (ora_tex_flag(mc, 0)) {
        ora_tex_slot(s, mc, mi, 0, sh->texCoord, t);
        m->baseColor = V3(t[0], t[1], t[2]);
        m->metalColor = m->baseColor;
    }
    if (ora_tex_flag(mc, 1)) {
        ora_tex_slot(s, mc, mi, 1, sh->texCoord, t);
        m->specularRoughness = t[0];
    }
    if (ora_tex_flag(mc, 2)) {
        ora_tex_slot(s, mc, mi, 2, sh->texCoord, t); /* unbound: 0 (InstanceData.glsl:147-153) */
        m->emissionColor = V3(t[0], t[1], t[2]);
    }
    if (ora_tex_flag(mc, 4)) {
        ora_tex_slot(s, mc, mi, 4, sh->texCoord, t);
        float N[3], T[3], n[3];
        v3store(N, sh->normal);
        v3store(T, sh->tangent);
        ora_normal_from_map(t, N, T, n);
        sh->normal = v3load(n);
        build_basis(sh->normal, &sh->tangent, &sh->bitangent);
    }
}

/* adjustRadiance, PathTracingCommon.slang:59-82 */
static v3 adjust_radiance(float maxLum, int displayErrors, v3 r)
{
    float lum = luminance(r);
    if (lum > maxLum) r = scale(r, maxLum / lum);
    if (isnan(r.x) || isnan(r.y) || isnan(r.z)) return displayErrors ? V3(INFINITY, 0.0f, 0.0f) : v3s(0.0f);
    if (isinf(r.x) || isinf(r.y) || isinf(r.z)) return displayErrors ? V3(0.0f, INFINITY, 0.0f) : v3s(0.0f);
    return r;
}

/* power heuristic, Sampling.slang:101-107 */
static float mis_weight(float f, float g) { return (f * f) / (f * f + g * g); }

/* buildBasis, Globals.slang:205-210 */
static void build_basis(v3 n, v3* t, v3* b)
{
    v3 bt = fabsf(n.y) < 0.999f ? V3(0.0f, 1.0f, 0.0f) : V3(1.0f, 0.0f, 0.0f);
    *t = normalize(cross(bt, n));
    *b = cross(n, *t);
}

/* sampleCone, Sampling.slang:72-86 */
static v3 sample_cone(const float r[2], v3 dir, float cosThetaMax)
{
    v3 xb, yb;
    build_basis(dir, &xb, &yb);
    float cosTheta = (1.0f - r[0]) + r[0] * cosThetaMax;
    float sinTheta = sqrtf(1.0f - cosTheta * cosTheta);
    float phi = r[1] * 2.0f * ORA_PI;
    return add(add(scale(xb, cosf(phi) * sinTheta), scale(yb, sinf(phi) * sinTheta)), scale(dir, cosTheta));
}

/* shadeDirectionalLight, ShadeFunctions.slang:51-89 */
static v3 shade_directional(const ora_distant_light* l, const ora_material* m, const ora_shading* sh, v3 V,
                            uint32_t rng[2], v3* shadowDir)
{
    float r[2];
    ora_random2d(rng, r);
    v3 L = sample_cone(r, v3load(l->direction), l->cos_radius);
    v3 f = ora_eval_material(m, sh, V, L);
    if (is_black(f)) {
        *shadowDir = v3s(0.0f);
        return v3s(0.0f);
    }
    v3 radiance = scale(v3load(l->color_intensity), l->color_intensity[3]);
    *shadowDir = L;
    return mul(radiance, f);
}

typedef struct {
    const ora_scene* s;
    const ora_material_constants* mats;
    uint32_t nmat;
    const ora_env* env;
    const ora_frame_params* fp;
    uint32_t W, H, x, y;
    ora_counters* cnt;
    const float* transmit; /* per-instance shadow transmittance, or NULL for forced-opaque shadows */
} pt_ctx;

/* The shadow any-hit shader's opacity per instance (HitShaderEntryPoints.slang:489-553):
 * xyz = 1 - opacity * (1 - transmission * transmissionColor); w = 1 for an opaque material
 * (opacity 1, no opacity image, no transmission: MaterialBase::isOpaque).  Constants only:
 * opacity images are not sampled for shadows. */
static float* make_transmit(const ora_scene* s, const ora_material_constants* mats, uint32_t nmat)
{
    float* t = (float*)malloc(sizeof(float) * 4 * (s->ninst ? s->ninst : 1));
    for (uint32_t i = 0; i < s->ninst; i++) {
        uint32_t mi = s->inst[i].material < nmat ? s->inst[i].material : 0;
        const ora_material_constants* m = &mats[mi];
        int opaque = m->opacity[0] == 1.0f && m->opacity[1] == 1.0f && m->opacity[2] == 1.0f &&
                     m->transmission == 0.0f && !ora_tex_flag(m, 3); /* the opacity slot */
        for (int k = 0; k < 3; k++)
            t[4 * i + k] = 1.0f - m->opacity[k] * (1.0f - m->transmission * m->transmissionColor[k]);
        t[4 * i + 3] = opaque ? 1.0f : 0.0f;
    }
    return t;
}

/* traceShadowRay, RayTrace.slang:221-270: occlusion only when shadows are forced opaque
 * (HGI forces them, HGIRenderer.cpp:490-491), otherwise visibility through non-opaque
 * instances. */
static v3 trace_shadow(const pt_ctx* c, v3 org, v3 dir)
{
    ora_ray r;
    v3store(r.org, org);
    v3store(r.dir, dir);
    r.tmin = ORA_RAY_TMIN;
    r.tmax = INFINITY;
    if (c->cnt) c->cnt->shadow_rays++;
    if (c->transmit) {
        float vis[3];
        ora_trace_transmit_one(c->s, c->transmit, &r, vis);
        return V3(vis[0], vis[1], vis[2]);
    }
    return v3s(ora_trace_any_one(c->s, &r) ? 0.0f : 1.0f);
}

typedef struct {
    uint32_t rng[2];
    float totalDistance;
    v3 rayContribution, nextRayOrigin, nextRayDirection, direct, indirect;
    int isPrimaryRay, pathEnded, skipContribution, rayType, nextRayType, misLobeID, rayDepth, isInside;
    float outputAlpha;
} payload;

static void env_eval(const pt_ctx* c, v3 dir, int asBg, v3* out)
{
    float d[3] = { dir.x, dir.y, dir.z }, o[3];
    ora_env_evaluate(c->env, d, asBg, c->x, c->y, c->W, c->H, o);
    *out = v3load(o);
}

/* shadeEnvironmentLightMIS, ShadeFunctions.slang:135-259 */
static void shade_env_mis(const pt_ctx* c, const ora_material* m, const ora_shading* sh, v3 V, uint32_t rng[2],
                          int* lobeID, int* emitsMat, v3* matResult, v3* matDir, int* emitsLight, v3* lightResult,
                          v3* lightDir)
{
    *emitsMat = 0;
    *emitsLight = 0;
    float r[2];
    ora_random2d(rng, r);
    v3 L;
    float matPDF;
    v3 f = ora_sample_material(m, sh, V, r, &L, &matPDF, lobeID);
    if (*lobeID == ORA_TRANSMISSION_LOBE) return;
    v3 Le;
    env_eval(c, L, *lobeID == ORA_TRANSMISSION_LOBE, &Le);
    if (!is_black(Le) && !is_black(f)) {
        float lightPDF = c->env->light_rgb ? luminance(Le) / c->env->light_integral : 0.25f * ORA_PI_INV;
        float w = mis_weight(matPDF, lightPDF);
        *matResult = divs(mul(scale(Le, w), f), matPDF);
        *matDir = L;
        *emitsMat = 1;
    }
    float lightPDF;
    float Lf[3], Lef[3];
    ora_env_sample(c->env, r, Lf, &lightPDF, Lef);
    L = v3load(Lf);
    Le = v3load(Lef);
    int evLobe;
    f = ora_eval_material_pdf(m, sh, V, L, r, &matPDF, &evLobe);
    if (!is_black(Le) && !is_black(f)) {
        float w = mis_weight(lightPDF, matPDF);
        *lightResult = divs(mul(scale(Le, w), f), lightPDF);
        *lightDir = L;
        *emitsLight = 1;
    }
}

/* setupIndirectLightRay, ShadeFunctions.slang:263-310 */
static int setup_indirect(const ora_material* m, const ora_shading* sh, v3 V, uint32_t rng[2], v3* dir, v3* org,
                          v3* color)
{
    float r[2];
    ora_random2d(rng, r);
    v3 L;
    float pdf;
    int lobe;
    v3 f = ora_sample_material(m, sh, V, r, &L, &pdf, &lobe);
    if (is_black(f)) return ORA_NO_SECONDARY_RAY;
    int isTrans = lobe == ORA_TRANSMISSION_LOBE;
    *org = (dot(V, sh->normal) < 0.0f || isTrans) ? sh->geomPosition : sh->position;
    *dir = L;
    *color = divs(f, pdf);
    return lobe;
}

/* INSTANCE_CLOSEST_HIT_SHADER, HitShaderEntryPoints.slang:35-486 (no layers, no ground plane). */
static void closest_hit(const pt_ctx* c, payload* pl, v3 rayDir, const ora_hit* hit)
{
    const ora_frame_params* fp = c->fp;
    v3 nextRayContribution = v3s(1.0f);
    pl->totalDistance += hit->t;
    v3 V = neg(rayDir);
    ora_shading sh;
    shading_data(c->s, hit->inst, hit->prim, hit->u, hit->v, &sh);
    /* Layer loop, HitShaderEntryPoints.slang:84-113 (DirectX semantics; HGI has no
     * layers and runs it once): layers from the top down to the base material, each
     * drawing its opacity coin (:235-240); a layer passed through scales the
     * throughput by transparency / P (:439). */
    const ora_scene* s = c->s;
    uint32_t lfirst = s->layer_off ? s->layer_off[hit->inst] : 0;
    uint32_t lcount = s->layer_off ? s->layer_off[hit->inst + 1] - lfirst : 0;
    ora_material m;
    float P = 0.0f;
    int hitLayer = 0;
    uint32_t hitMat = 0;
    float r[2];
    /* computeLayerUV (GlobalBufferAccessors.slang:116-127): a layer with geometry of its own
     * reads that geometry's texture coordinates at the hit's vertices; the base layer gets
     * the instance's back (HitShaderEntryPoints.slang:98-111). */
    const float baseUV[2] = { sh.texCoord[0], sh.texCoord[1] };
    const uint32_t* tri = s->meshes[s->inst[hit->inst].mesh].idx + 3 * (size_t)hit->prim;
    const float bw[3] = { 1.0f - hit->u - hit->v, hit->u, hit->v };
    for (int li = (int)lcount; li >= 0; li--) {
        uint32_t mi = li == 0 ? s->inst[hit->inst].material : s->layer_mat[lfirst + (uint32_t)li - 1];
        const float* luv = (li > 0 && s->layer_uv) ? s->layer_uv[lfirst + (uint32_t)li - 1] : NULL;
        sh.texCoord[0] = baseUV[0];
        sh.texCoord[1] = baseUV[1];
        if (luv) {
            sh.texCoord[0] = sh.texCoord[1] = 0.0f;
            for (int k = 0; k < 3; k++) {
                sh.texCoord[0] = sh.texCoord[0] + luv[2 * (size_t)tri[k]] * bw[k];
                sh.texCoord[1] = sh.texCoord[1] + luv[2 * (size_t)tri[k] + 1] * bw[k];
            }
        }
        if (mi >= c->nmat) mi = 0;
        ora_material_from_constants(&c->mats[mi], &m);
        if (ora_tex_flag(&c->mats[mi], 3)) { /* opacity texture */
            float t[3];
            ora_tex_slot(s, &c->mats[mi], mi, 3, sh.texCoord, t);
            m.opacity = V3(t[0], t[1], t[2]);
        }
        hitMat = mi;
        v3 transparency = sub(v3s(1.0f), m.opacity);
        P = luminance(transparency);
        ora_random2d(pl->rng, r);
        hitLayer = r[0] >= P;
        if (hitLayer) break;
        pl->rayContribution = mul(pl->rayContribution, divs(transparency, P));
    }
    v3 dirShadowDir = v3s(0.0f), dirLightColor = v3s(0.0f);
    int hasDirLight = 0;
    if (hitLayer) {
        material_textures(s, &c->mats[hitMat], hitMat, &m, &sh);
        /* shadeEmission, ShadeFunctions.slang:31-48 */
        v3 ecb = lerp3(v3s(1.0f), m.coatColor, m.coat);
        pl->direct = add(pl->direct, mul(scale(m.emissionColor, m.emission), ecb));

        int evalAll = fp->light_count > 1;
        if (evalAll) {
            v3 allOrigin = dot(V, sh.normal) < 0.0f ? sh.geomPosition : sh.position;
            for (int li = 0; li < fp->light_count; li++) {
                if (fp->lights[li].color_intensity[3] <= 0.0f) continue;
                v3 sd;
                v3 lr = shade_directional(&fp->lights[li], &m, &sh, V, pl->rng, &sd);
                if (is_black(lr)) continue;
                v3 vis = trace_shadow(c, allOrigin, sd);
                pl->direct = add(pl->direct, mul(vis, lr));
            }
        } else if (fp->light_count > 0) {
            float weights[4] = { 0.0f, 0.0f, 0.0f, 0.0f };
            float total = 0.0f;
            for (int i = 0; i < fp->light_count; i++) {
                float w = luminance(v3load(fp->lights[i].color_intensity)) * fp->lights[i].color_intensity[3];
                weights[i] = w;
                total += w;
            }
            int li = 0;
            float selPdf = 1.0f / (float)fp->light_count;
            if (total > 0.0f) {
                ora_random2d(pl->rng, r);
                float rr = r[0] * total;
                float cum = 0.0f;
                for (int i = 0; i < fp->light_count; i++) {
                    cum += weights[i];
                    if (rr <= cum || i == fp->light_count - 1) {
                        li = i;
                        selPdf = weights[i] / total;
                        break;
                    }
                }
            } else {
                ora_random2d(pl->rng, r);
                li = (int)(r[0] * (float)fp->light_count);
            }
            if (fp->lights[li].color_intensity[3] > 0.0f) {
                dirLightColor = shade_directional(&fp->lights[li], &m, &sh, V, pl->rng, &dirShadowDir);
                dirLightColor = divs(dirLightColor, maxf(selPdf, 1e-4f));
                hasDirLight = 1;
            }
        }

        int misLobe, emitsMat, emitsLight;
        v3 matRes = v3s(0.0f), matDir = v3s(0.0f), lightRes = v3s(0.0f), lightDir = v3s(0.0f);
        shade_env_mis(c, &m, &sh, V, pl->rng, &misLobe, &emitsMat, &matRes, &matDir, &emitsLight, &lightRes,
                      &lightDir);

        pl->rayContribution = mul(pl->rayContribution, divs(m.opacity, 1.0f - P));
        pl->outputAlpha = (pl->outputAlpha == 0.0f && pl->rayType == ORA_TRANSMISSION_LOBE) ? 0.0f : 1.0f;
        pl->nextRayType = setup_indirect(&m, &sh, V, pl->rng, &pl->nextRayDirection, &pl->nextRayOrigin,
                                         &nextRayContribution);
        v3 shadowOrigin = dot(V, sh.normal) < 0.0f ? sh.geomPosition : sh.position;
        if (hasDirLight) {
            v3 vis = trace_shadow(c, shadowOrigin, dirShadowDir);
            pl->direct = add(pl->direct, mul(vis, dirLightColor));
        }
        if (emitsLight || emitsMat) {
            v3 misResult = v3s(0.0f);
            if (emitsMat) {
                v3 vis = trace_shadow(c, shadowOrigin, matDir);
                misResult = add(misResult, mul(vis, matRes));
            }
            if (emitsLight) {
                v3 vis = trace_shadow(c, shadowOrigin, lightDir);
                misResult = add(misResult, mul(vis, lightRes));
            }
            pl->misLobeID = misLobe;
            pl->indirect = misResult;
        }
        if (pl->nextRayType == ORA_NO_SECONDARY_RAY) pl->pathEnded = 1;
        pl->isPrimaryRay = 0;
    }

    if (!hitLayer) {
        pl->nextRayDirection = neg(V);
        pl->nextRayOrigin = sh.geomPosition;
        pl->skipContribution = 1;
        return;
    }
    if (pl->nextRayType == ORA_TRANSMISSION_LOBE) pl->isInside = !pl->isInside;
    pl->direct = mul(pl->direct, pl->rayContribution);
    pl->indirect = mul(pl->indirect, pl->rayContribution);
    pl->rayContribution = mul(pl->rayContribution, nextRayContribution);
    if (fp->rr_enabled && pl->rayDepth + 1 >= fp->rr_start_depth) {
        v3 tp = pl->rayContribution;
        float surv = clampf(maxf(tp.x, maxf(tp.y, tp.z)), 0.05f, 1.0f);
        ora_random2d(pl->rng, r);
        if (r[0] > surv) pl->pathEnded = 1;
        else pl->rayContribution = divs(pl->rayContribution, surv);
    }
}

/* InstanceMissShader, MainEntryPoints.slang:354-426 (its ground plane step runs before it, in
 * pixel_sample). */
static void miss(const pt_ctx* c, payload* pl, v3 rayDir)
{
    int shadeBg = pl->isPrimaryRay || pl->rayType == ORA_TRANSMISSION_LOBE ||
                  pl->rayType == ORA_GROUND_PLANE_REFLECTION_RAY;
    if (shadeBg) {
        v3 e;
        env_eval(c, rayDir, 1, &e);
        pl->direct = mul(e, pl->rayContribution);
    }
    pl->pathEnded = 1;
}

/* ---- ground plane (DirectX only: Shaders/GroundPlane.slang:24-290) ---- */

void ora_ggx_smith_sample(const float V[3], const float f0[3], float roughness, const float rnd[2], float L[3],
                          float* pdf, float out[3])
{
    v3 l;
    v3 f = ora_ggx_smith_brdf_sample(v3load(V), v3load(f0), roughness, rnd, &l, pdf);
    v3store(L, l);
    v3store(out, f);
}

int ora_ground_plane_hit(const ora_ground_plane* gp, const float org[3], const float dir[3], float pos[3],
                         float* dist)
{
    v3 n = v3load(gp->normal), o = v3load(org), d = v3load(dir);
    float pd = dot(d, n);
    if (pd >= 0.0f) return 0;
    float t = dot(sub(v3load(gp->position), o), n) / pd;
    if (t < 0.0f) return 0;
    *dist = t;
    v3store(pos, add(o, scale(d, t)));
    return 1;
}

/* evaluateGroundPlane (:232-290) for a primary ray whose closest hit is at tHit (INFINITY
 * on a miss).  A reflection sample ends this segment ("absorbed", returns 1: the path
 * continues from the plane as a ground-plane reflection ray); otherwise the contribution
 * is tinted by the plane (transparency, matte shadow) and the hit or miss is shaded. */
static int ground_plane(const pt_ctx* c, payload* pl, v3 org, v3 dir, float tHit)
{
    const ora_scene* s = c->s;
    const ora_ground_plane* g = &s->ground;
    float o[3], d[3], pos[3], dist;
    v3store(o, org);
    v3store(d, dir);
    if (!ora_ground_plane_hit(g, o, d, pos, &dist) || !(dist < tHit)) return 0;
    v3 N = v3load(g->normal), T = v3load(s->ground_t), B = v3load(s->ground_b), P = v3load(pos);
    v3 color = v3s(1.0f);
    float r[2];
    if (g->reflection_opacity > 0.0f) {
        v3 transparency = v3s(1.0f - g->reflection_opacity);
        float Pr = luminance(transparency);
        ora_random2d(pl->rng, r);
        if (r[0] >= Pr) {
            /* setupMatteReflectionIndirectLightRay (:24-65) */
            v3 V = V3(dot(neg(dir), T), dot(neg(dir), B), dot(neg(dir), N));
            v3 L;
            float pdf;
            ora_random2d(pl->rng, r);
            v3 bc = ora_ggx_smith_brdf_sample(V, v3load(g->reflection_color), g->reflection_roughness, r, &L, &pdf);
            bc = scale(bc, fabsf(L.z));
            int reflect = !is_black(bc);
            if (reflect) color = divs(bc, pdf);
            color = scale(color, g->reflection_opacity / (1.0f - Pr));
            pl->rayContribution = mul(pl->rayContribution, color);
            if (!reflect) return 0; /* NO_SECONDARY_RAY: on to the hit, tinted, no shadow */
            pl->nextRayType = ORA_GROUND_PLANE_REFLECTION_RAY;
            pl->nextRayOrigin = P;
            pl->nextRayDirection = add(add(scale(T, L.x), scale(B, L.y)), scale(N, L.z));
            pl->isPrimaryRay = 0;
            return 1;
        }
        color = divs(transparency, Pr);
    }
    if (g->shadow_opacity > 0.0f) {
        /* shadeMatteShadow (:67-101): one environment sample, an opaque shadow ray */
        float Lf[3], pdf, Le[3];
        ora_random2d(pl->rng, r);
        ora_env_sample(c->env, r, Lf, &pdf, Le);
        v3 L = v3load(Lf);
        if (!(dot(L, N) < 0.0f)) {
            ora_ray sr;
            v3store(sr.org, P);
            v3store(sr.dir, L);
            sr.tmin = ORA_RAY_TMIN;
            sr.tmax = INFINITY;
            if (c->cnt) c->cnt->shadow_rays++;
            v3 vis = v3s(ora_trace_any_one(s, &sr) ? 0.0f : 1.0f);
            float a = luminance(sub(v3s(1.0f), vis)) * g->shadow_opacity;
            color = mul(color, lerp3(v3s(1.0f), v3load(g->shadow_color), a));
        }
    }
    pl->rayContribution = mul(pl->rayContribution, color);
    return 0;
}

/* computeCameraRay, RayTrace.slang:130-182 */
static void camera_ray(const ora_frame_params* fp, uint32_t x, uint32_t y, uint32_t W, uint32_t H,
                       uint32_t rng[2], v3* org, v3* dir)
{
    float sc[2] = { (float)x, (float)y };
    if (fp->temporal_jitter) {
        sc[0] += 0.5f + fp->jitter[0];
        sc[1] += 0.5f + fp->jitter[1];
    } else {
        float r[2];
        ora_random2d(rng, r);
        sc[0] += r[0];
        sc[1] += r[1];
    }
    float ndc[2] = { (sc[0] / (float)W) * 2.0f - 1.0f, (sc[1] / (float)H) * 2.0f - 1.0f };
    ndc[1] = -ndc[1];
    v3 right = v3load(fp->inv_view + 0), up = v3load(fp->inv_view + 4), front = neg(v3load(fp->inv_view + 8));
    float size[2] = { fp->view_size[0] * 0.5f, fp->view_size[1] * 0.5f };
    v3 off = add(scale(right, size[0] * ndc[0]), scale(up, size[1] * ndc[1]));
    v3 eye = v3load(fp->inv_view + 12);
    if (fp->is_ortho) {
        *dir = front;
        *org = add(eye, off);
    } else {
        *dir = normalize(add(front, off));
        *org = eye;
    }
    if (fp->lens_radius > 0.0f) {
        v3 focal = add(*org, scale(*dir, fp->focal_distance));
        float r[2];
        ora_random2d(rng, r);
        float rr = fp->lens_radius * sqrtf(r[0]);
        float phi = 2.0f * ORA_PI * r[1];
        float ox = rr * cosf(phi), oy = rr * sinf(phi);
        *org = add(add(*org, scale(right, ox)), scale(up, oy));
        *dir = normalize(sub(focal, *org));
    }
}

/* The camera ray of one pixel-sample, with the path's RNG state (sample_index
 * includes the seed offset), for unit tests of computeCameraRay. */
void ora_camera_ray(const ora_frame_params* fp, uint32_t W, uint32_t H, uint32_t x, uint32_t y,
                    uint32_t sample_index, float org[3], float dir[3])
{
    uint32_t rng[2];
    ora_pcg2d_init(sample_index, W, x, y, rng);
    v3 o, d;
    camera_ray(fp, x, y, W, H, rng, &o, &d);
    v3store(org, o);
    v3store(dir, d);
}

/* RayGenShader, MainEntryPoints.slang:70-334 (HGI: no AOVs, no denoising). */
static void pixel_sample(const ora_scene* s, const ora_material_constants* mats, uint32_t nmat, const ora_env* env,
                         const ora_frame_params* fp, uint32_t W, uint32_t H, uint32_t x, uint32_t y,
                         uint32_t sample_index, float out[4], ora_counters* cnt, const float* transmit)
{
    pt_ctx c = { s, mats, nmat, env, fp, W, H, x, y, cnt, transmit };
    int maxDepth = fp->trace_depth;
    payload pl;
    memset(&pl, 0, sizeof(pl));
    ora_pcg2d_init(sample_index, W, x, y, pl.rng);
    pl.rayContribution = v3s(1.0f);
    pl.rayType = ORA_CAMERA_RAY;
    pl.misLobeID = ORA_NO_SECONDARY_RAY;
    pl.outputAlpha = 0.0f;
    pl.isPrimaryRay = 1;
    v3 rayOrg, rayDir;
    camera_ray(fp, x, y, W, H, pl.rng, &rayOrg, &rayDir);
    v3 outDirect = v3s(0.0f), outDiffuse = v3s(0.0f), outGlossy = v3s(0.0f);
    float outAlpha = 0.0f;
    for (int depth = 0; depth <= maxDepth && !pl.pathEnded; depth++) {
        if (depth == maxDepth) {
            int shadeBg = pl.isPrimaryRay || pl.rayType == ORA_TRANSMISSION_LOBE ||
                          pl.rayType == ORA_GROUND_PLANE_REFLECTION_RAY;
            v3 e;
            env_eval(&c, rayDir, shadeBg, &e);
            pl.direct = mul(e, pl.rayContribution);
            pl.pathEnded = 1;
            /* pl.indirect keeps the previous segment's value (not cleared here) */
        } else {
            float minT = depth == 0 ? 0.0f : ORA_RAY_TMIN;
            pl.rayDepth = depth;
            pl.direct = v3s(0.0f);
            pl.indirect = v3s(0.0f);
            ora_ray r;
            v3store(r.org, rayOrg);
            v3store(r.dir, rayDir);
            r.tmin = minT;
            r.tmax = INFINITY;
            ora_hit hit;
            if (cnt) cnt->extension_rays++;
            ora_trace_closest_one(s, &r, &hit);
            int absorbed = 0;
            if (pl.isPrimaryRay && s->ground.enabled)
                absorbed = ground_plane(&c, &pl, rayOrg, rayDir, hit.inst != 0xFFFFFFFFu ? hit.t : INFINITY);
            if (absorbed) {
                /* the segment ends at the plane: nothing shaded, no roulette */
            } else if (hit.inst != 0xFFFFFFFFu) {
                closest_hit(&c, &pl, rayDir, &hit);
            } else {
                miss(&c, &pl, rayDir);
            }
            rayOrg = pl.nextRayOrigin;
            rayDir = pl.nextRayDirection;
            outAlpha = pl.outputAlpha;
            if (pl.skipContribution) {
                pl.skipContribution = 0;
                continue;
            }
        }
        if (depth == 0) {
            pl.isPrimaryRay = 0;
            outDirect = pl.direct;
            if (pl.misLobeID == ORA_DIFFUSE_LOBE) outDiffuse = pl.indirect;
            else outGlossy = pl.indirect;
        } else {
            v3 radiance = add(pl.direct, pl.indirect);
            if (pl.rayType == ORA_DIFFUSE_LOBE) outDiffuse = add(outDiffuse, radiance);
            else outGlossy = add(outGlossy, radiance);
        }
        pl.rayType = pl.nextRayType;
    }
    v3 fin = add(add(outDirect, outDiffuse), outGlossy);
    fin = adjust_radiance(fp->max_luminance, fp->display_errors, fin);
    out[0] = fin.x;
    out[1] = fin.y;
    out[2] = fin.z;
    out[3] = outAlpha;
}

void ora_pixel_sample(const ora_scene* s, const ora_material_constants* mats, uint32_t nmat, const ora_env* env,
                      const ora_frame_params* fp, uint32_t W, uint32_t H, uint32_t x, uint32_t y,
                      uint32_t sample_index, float out[4], ora_counters* cnt)
{
    float* transmit = fp->force_opaque_shadows ? NULL : make_transmit(s, mats, nmat);
    pixel_sample(s, mats, nmat, env, fp, W, H, x, y, sample_index, out, cnt, transmit);
    free(transmit);
}

typedef struct {
    const ora_scene* s; const ora_material_constants* mats; uint32_t nmat; const ora_env* env;
    const ora_frame_params* fp; uint32_t W, H, start, count; const uint32_t* pixels; uint32_t npix;
    float* accum; ora_counters cnt; int tid, nthreads; const float* transmit;
} render_job;

static void* render_worker(void* p)
{
    render_job* j = (render_job*)p;
    uint32_t total = j->npix ? j->npix : j->W * j->H;
    for (uint32_t k = (uint32_t)j->tid; k < total; k += (uint32_t)j->nthreads) {
        uint32_t pix = j->npix ? j->pixels[k] : k;
        uint32_t x = pix % j->W, y = pix / j->W;
        float* a = j->accum + 4 * (size_t)pix;
        for (uint32_t si = j->start; si < j->start + j->count; si++) {
            float res[4];
            pixel_sample(j->s, j->mats, j->nmat, j->env, j->fp, j->W, j->H, x, y, si, res, &j->cnt, j->transmit);
            if (si > 0) {
                /* Accumulation.glsl:22-30: mix(prev, new, 1/(i+1)) */
                float t = 1.0f / (float)(si + 1);
                for (int ch = 0; ch < 4; ch++) res[ch] = a[ch] * (1.0f - t) + res[ch] * t;
            }
            for (int ch = 0; ch < 4; ch++) a[ch] = res[ch];
        }
    }
    return NULL;
}

void ora_render(const ora_scene* s, const ora_material_constants* mats, uint32_t nmat, const ora_env* env,
                const ora_frame_params* fp, uint32_t W, uint32_t H, uint32_t start, uint32_t count,
                const uint32_t* pixels, uint32_t npix, int nthreads, float* accum, ora_counters* counters)
{
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    int started[256];
    render_job* jobs = (render_job*)calloc((size_t)nthreads, sizeof(render_job));
    float* transmit = fp->force_opaque_shadows ? NULL : make_transmit(s, mats, nmat);
    /* explicit stack size: the worker must not depend on the process's RLIMIT_STACK */
    pthread_attr_t attr;
    pthread_attr_init(&attr);
    pthread_attr_setstacksize(&attr, (size_t)16 << 20);
    for (int t = 0; t < nthreads; t++) {
        jobs[t] = (render_job){ s, mats, nmat, env, fp, W, H, start, count, pixels, npix, accum, { 0, 0 }, t,
                                nthreads, transmit };
        started[t] = pthread_create(&th[t], &attr, render_worker, &jobs[t]) == 0;
        if (!started[t]) render_worker(&jobs[t]); /* no thread available: run this share inline */
    }
    pthread_attr_destroy(&attr);
    ora_counters tot = { 0, 0 };
    for (int t = 0; t < nthreads; t++) {
        if (started[t]) pthread_join(th[t], NULL);
        tot.extension_rays += jobs[t].cnt.extension_rays;
        tot.shadow_rays += jobs[t].cnt.shadow_rays;
    }
    free(jobs);
    free(transmit);
    if (counters) *counters = tot;
}

/* Texture coordinate of a hit (Geometry.slang:199-203), for tests. */
void ora_hit_uv(const ora_scene* s, uint32_t inst, uint32_t prim, float bu, float bv, float out[2])
{
    ora_shading sh;
    shading_data(s, inst, prim, bu, bv, &sh);
    out[0] = sh.texCoord[0];
    out[1] = sh.texCoord[1];
}

/* The AOV pass: guides of the first segment of one camera ray, as the primary
 * closest-hit / miss shaders write them (HitShaderEntryPoints.slang:184-214,
 * MainEntryPoints.slang:391-405; motion vectors GlobalPipelineState.slang:126-172). */
void ora_aov_pixel(const ora_scene* s, const ora_material_constants* mats, uint32_t nmat,
                   const ora_frame_params* fp, uint32_t W, uint32_t H, uint32_t x, uint32_t y,
                   uint32_t sample_index, const float view_proj[16], const float view_proj_prev[16],
                   const float inv_view_prev[16], float sky_view_z, float out[13])
{
    uint32_t rng[2];
    ora_pcg2d_init(sample_index, W, x, y, rng);
    v3 org, dir;
    camera_ray(fp, x, y, W, H, rng, &org, &dir);
    ora_ray r;
    v3store(r.org, org);
    v3store(r.dir, dir);
    r.tmin = 0.0f;
    r.tmax = INFINITY;
    ora_hit h;
    ora_trace_closest_one(s, &r, &h);
    if (h.inst == 0xFFFFFFFFu) {
        float d4[4] = { dir.x, dir.y, dir.z, 0.0f }, mv[2];
        ora_screen_motion(view_proj, view_proj_prev, d4, W, H, mv);
        float o[13] = { sky_view_z, 0, 0, 0, 0.1f, 0, 0, 0, 0.1f, mv[0], mv[1], 0, 0 };
        memcpy(out, o, sizeof(o));
        return;
    }
    float pos[3] = { org.x + dir.x * h.t, org.y + dir.y * h.t, org.z + dir.z * h.t };
    float vz = ora_view_z(fp->inv_view, pos);
    ora_shading sh;
    shading_data(s, h.inst, h.prim, h.u, h.v, &sh);
    uint32_t mi = s->inst[h.inst].material;
    if (mi >= nmat) mi = 0;
    ora_material m;
    ora_material_from_constants(&mats[mi], &m);
    material_textures(s, &mats[mi], mi, &m, &sh); /* base colour, roughness, normal map */
    v3 n = normalize(sh.normal);
    float p4[4] = { pos[0], pos[1], pos[2], 1.0f }, mv[2];
    ora_screen_motion(view_proj, view_proj_prev, p4, W, H, mv);
    float prevZ = ora_view_z(inv_view_prev, pos);
    float rough = m.specularRoughness > 0.05f ? m.specularRoughness : 0.05f;
    float o[13] = { vz, n.x, n.y, n.z, rough, m.baseColor.x, m.baseColor.y, m.baseColor.z, m.metalness,
                    mv[0], mv[1], vz - prevZ, 0.0f };
    memcpy(out, o, sizeof(o));
}
