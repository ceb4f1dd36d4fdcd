/* Oracle: material textures (test infrastructure only).
 * Restates the textured part of evaluateDefaultMaterial (Shaders/DefaultMaterial.slang:
 * applyUVTransform :97-104 with rotateUV :86-93, calculateNormalFromMap :107-127,
 * the texture branches :171-233) and the sampling the HGI backend configures
 * (HGI/HGISampler.cpp:22-60: bilinear, one mip level per HGI/HGIImage.cpp:35, address
 * mode per sampler, opaque-white border; HGIImage.cpp:98-100: RGBA8 images with
 * linearize are sRGB-decoded per texel before filtering). */
#include "ora.h"
#include "ora_geom.h"
#include "ora_math.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* sRGB transfer function undone (IEC 61966-2-1). */
float ora_srgb_to_linear(float c)
{
    return c <= 0.04045f ? c / 12.92f : powf((c + 0.055f) / 1.055f, 2.4f);
}

/* xf = offset.xy, scale.xy, pivot.xy, rotation (degrees) */
void ora_uv_transform(const float uv[2], const float xf[7], float out[2])
{
    float sx = (uv[0] - xf[4]) / xf[2], sy = (uv[1] - xf[5]) / xf[3];
    float r = xf[6] * (ORA_PI / 180.0f);
    float sa = sinf(r), ca = cosf(r);
    float rx = ca * sx + sa * sy, ry = -sa * sx + ca * sy;
    out[0] = (rx - xf[0]) + xf[4];
    out[1] = (ry - xf[1]) + xf[5];
}

/* Vulkan address modes on an integer texel coordinate. */
static int address(int i, int n, int mode, int* border)
{
    switch (mode) {
    case 1: { /* mirrored repeat */
        int t = i % (2 * n);
        if (t < 0) t += 2 * n;
        return t < n ? t : 2 * n - 1 - t;
    }
    case 2: return i < 0 ? 0 : (i >= n ? n - 1 : i); /* clamp to edge */
    case 3: { /* mirror once, then clamp */
        int j = i < 0 ? -1 - i : i;
        return j >= n ? n - 1 : j;
    }
    case 4: /* clamp to border */
        if (i < 0 || i >= n) *border = 1;
        return i < 0 ? 0 : (i >= n ? n - 1 : i);
    default: { /* repeat */
        int t = i % n;
        return t < 0 ? t + n : t;
    }
    }
}

void ora_texture_sample(const ora_texture* t, const float uv_in[2], float out[4])
{
    float u = isfinite(uv_in[0]) ? uv_in[0] : 0.0f, v = isfinite(uv_in[1]) ? uv_in[1] : 0.0f;
    float x = u * (float)t->w - 0.5f, y = v * (float)t->h - 0.5f;
    float x0f = floorf(x), y0f = floorf(y);
    float fx = x - x0f, fy = y - y0f;
    x0f = fminf(fmaxf(x0f, -1073741824.0f), 1073741824.0f);
    y0f = fminf(fmaxf(y0f, -1073741824.0f), 1073741824.0f);
    int x0 = (int)x0f, y0 = (int)y0f;
    float c[4][4];
    for (int k = 0; k < 4; k++) {
        int border = 0;
        int xi = address(x0 + (k & 1), (int)t->w, t->addr_u, &border);
        int yi = address(y0 + (k >> 1), (int)t->h, t->addr_v, &border);
        for (int ch = 0; ch < 4; ch++)
            c[k][ch] = border ? 1.0f : t->rgba[4 * ((size_t)yi * t->w + (size_t)xi) + ch];
    }
    for (int ch = 0; ch < 4; ch++) {
        float a = c[0][ch] + (c[1][ch] - c[0][ch]) * fx;
        float b = c[2][ch] + (c[3][ch] - c[2][ch]) * fx;
        out[ch] = a + (b - a) * fy;
    }
}

/* calculateNormalFromMap, DefaultMaterial.slang:107-127 (tangent space, scale 1). */
void ora_normal_from_map(const float texel[3], const float N[3], const float T[3], float out[3])
{
    v3 v = sub(scale(V3(texel[0], texel[1], texel[2]), 2.0f), v3s(1.0f));
    v3 n = v3load(N), t = v3load(T);
    v3 b = normalize(cross(n, t));
    v3store(out, normalize(add(add(scale(t, v.x), scale(b, v.y)), scale(n, v.z))));
}

/* Word offsets (from byte 224) of each slot's flag, offset, scale, pivot and rotation
 * in MaterialConstants (DefaultMaterialUniformBuffer.slang:43-71). */
static const int kWord[ORA_TEX_SLOTS][5] = {
    {0, 1, 4, 6, 8}, {9, 10, 12, 14, 16}, {17, 18, 20, 22, 24}, {25, 26, 28, 30, 32}, {33, 34, 36, 38, 40}};

int ora_tex_flag(const ora_material_constants* mc, int slot) { return mc->tex_block[kWord[slot][0]] != 0; }

static float word_f(const ora_material_constants* mc, int w)
{
    float f;
    memcpy(&f, &mc->tex_block[w], 4);
    return f;
}

int ora_tex_slot(const ora_scene* s, const ora_material_constants* mc, uint32_t mi, int slot, const float uv[2],
                 float out[3])
{
    out[0] = out[1] = out[2] = 0.0f;
    int ti = (s->slot_tex && mi < s->nmat_tex) ? s->slot_tex[(size_t)mi * ORA_TEX_SLOTS + slot] : -1;
    if (ti < 0 || (uint32_t)ti >= s->ntex) return 0;
    const int* w = kWord[slot];
    float xf[7] = {word_f(mc, w[1]), word_f(mc, w[1] + 1), word_f(mc, w[2]), word_f(mc, w[2] + 1),
                   word_f(mc, w[3]), word_f(mc, w[3] + 1), word_f(mc, w[4])};
    float tuv[2], c[4];
    ora_uv_transform(uv, xf, tuv);
    ora_texture_sample(&s->tex[ti], tuv, c);
    out[0] = c[0];
    out[1] = c[1];
    out[2] = c[2];
    return 1;
}

void ora_scene_set_textures(ora_scene* s, const ora_texture* tex, uint32_t ntex, const int32_t* slot_tex,
                            uint32_t nmat)
{
    for (uint32_t i = 0; i < s->ntex; i++) free((void*)s->tex[i].rgba);
    free(s->tex);
    free(s->slot_tex);
    s->tex = (ora_texture*)calloc(ntex ? ntex : 1, sizeof(ora_texture));
    for (uint32_t i = 0; i < ntex; i++) {
        size_t n = (size_t)tex[i].w * tex[i].h * 4;
        float* px = (float*)malloc(n * sizeof(float));
        memcpy(px, tex[i].rgba, n * sizeof(float));
        s->tex[i] = tex[i];
        s->tex[i].rgba = px;
    }
    s->ntex = ntex;
    s->slot_tex = (int32_t*)malloc(sizeof(int32_t) * ORA_TEX_SLOTS * (nmat ? nmat : 1));
    memcpy(s->slot_tex, slot_tex, sizeof(int32_t) * ORA_TEX_SLOTS * nmat);
    s->nmat_tex = nmat;
}
