/* Oracle: RNG, transforms, environment and alias map (test infrastructure only). */
#include "ora.h"
#include "ora_math.h"

#include <stdlib.h>
#include <string.h>

/* pcg2DInit, Random.slang:96-108: state = (pixel index, sample index). */
void ora_pcg2d_init(uint32_t sample_index, uint32_t width, uint32_t x, uint32_t y, uint32_t state[2])
{
    state[0] = y * width + x;
    state[1] = sample_index;
}

/* pcg2D, Random.slang:113-128 (Jarzynski & Olano, JCGT 9(3) 2020).  The state
 * advances by the LCG and the *output* is the PCG2D mix of the previous state;
 * pcg2DNext2D stores that output back as the new state (Random.slang:133). */
void ora_pcg2d(uint32_t state[2])
{
    uint32_t vx = state[0], vy = state[1];
    vx += vy * 1664525u;
    vy += vx * 1664525u;
    vx ^= vx >> 16u;
    vy ^= vy >> 16u;
    vx += vy * 1664525u;
    vy += vx * 1664525u;
    vx ^= vx >> 16u;
    vy ^= vy >> 16u;
    state[0] = vx;
    state[1] = vy;
}

/* pcg2DNext2D, Random.slang:131-137: u32 / float(0xFFFFFFFF).  float(0xFFFFFFFF)
 * rounds to 2^32, and the u32 itself is rounded to float first, so the result
 * lies in [0, 1] inclusive. */
void ora_random2d(uint32_t state[2], float out[2])
{
    ora_pcg2d(state);
    const float denom = (float)0xFFFFFFFFu;
    out[0] = (float)state[0] / denom;
    out[1] = (float)state[1] / denom;
}

void ora_random_stream(uint32_t sx, uint32_t sy, uint32_t n, float* out)
{
    uint32_t st[2] = { sx, sy };
    for (uint32_t i = 0; i < n; i++)
        ora_random2d(st, out + 2 * i);
}

/* Inverse of a row-major 3x4 affine transform: double-precision cofactors,
 * fixed operation order, rounded to float. */
int ora_inverse3x4(const float mf[12], float out[12])
{
    double m[12];
    for (int i = 0; i < 12; i++) m[i] = (double)mf[i];
    double a = m[0], b = m[1], c = m[2], d = m[4], e = m[5], f = m[6], g = m[8], h = m[9], k = m[10];
    double A = e * k - f * h;
    double B = f * g - d * k;
    double C = d * h - e * g;
    double det = a * A + b * B + c * C;
    if (det == 0.0) {
        memset(out, 0, 12 * sizeof(float));
        return -1;
    }
    double r[9];
    r[0] = A / det;
    r[1] = (c * h - b * k) / det;
    r[2] = (b * f - c * e) / det;
    r[3] = B / det;
    r[4] = (a * k - c * g) / det;
    r[5] = (c * d - a * f) / det;
    r[6] = C / det;
    r[7] = (b * g - a * h) / det;
    r[8] = (a * e - b * d) / det;
    double tx = m[3], ty = m[7], tz = m[11];
    for (int i = 0; i < 3; i++) {
        out[4 * i + 0] = (float)r[3 * i + 0];
        out[4 * i + 1] = (float)r[3 * i + 1];
        out[4 * i + 2] = (float)r[3 * i + 2];
        out[4 * i + 3] = (float)(-((r[3 * i + 0] * tx + r[3 * i + 1] * ty) + r[3 * i + 2] * tz));
    }
    return 0;
}

/* ---------------- environment ---------------- */

/* directionToLatLongUV, Environment.slang:35-42 */
void ora_dir_to_latlong(const float dir[3], float uv[2])
{
    uv[0] = atan2f(dir[0], -dir[2]) * ORA_PI_INV * 0.5f + 0.5f;
    uv[1] = -asinf(dir[1]) * ORA_PI_INV + 0.5f;
}

/* latLongUVToDirection, Environment.slang:46-59 */
void ora_latlong_to_dir(const float uv[2], float dir[3])
{
    float phi = (uv[0] + 0.25f) * 2.0f * ORA_PI;
    float theta = uv[1] * ORA_PI;
    float sp = sinf(phi), cp = cosf(phi), st = sinf(theta), ct = cosf(theta);
    dir[0] = cp * st;
    dir[1] = ct;
    dir[2] = sp * st;
}

/* Texture lookup with repeat addressing.  filter 0 = nearest, 1 = bilinear.
 * A non-finite coordinate reads as 0: a NaN direction reaches the lookup when
 * sampleHemisphere (Sampling.slang:39-52) normalizes n + d with d = -n, which the
 * u32-to-float mapping makes possible (a draw of exactly 1.0).  Shader samplers
 * leave NaN coordinates implementation-defined; gfx950's float-to-int
 * conversion gives 0, and both paths state that explicitly. */
static v3 tex_lookup(const float* rgb, uint32_t w, uint32_t h, float u, float v, int filter)
{
    if (!isfinite(u)) u = 0.0f;
    if (!isfinite(v)) v = 0.0f;
    if (filter == 0) {
        float fu = u - floorf(u), fv = v - floorf(v);
        int ix = (int)(fu * (float)w), iy = (int)(fv * (float)h);
        if (ix >= (int)w) ix = (int)w - 1;
        if (iy >= (int)h) iy = (int)h - 1;
        return v3load(rgb + 3 * ((size_t)iy * w + (size_t)ix));
    }
    float x = u * (float)w - 0.5f, y = v * (float)h - 0.5f;
    float x0f = floorf(x), y0f = floorf(y);
    float fx = x - x0f, fy = y - y0f;
    int x0 = (int)x0f, y0 = (int)y0f;
    int xs[2] = { x0, x0 + 1 }, ys[2] = { y0, y0 + 1 };
    for (int i = 0; i < 2; i++) {
        xs[i] %= (int)w; if (xs[i] < 0) xs[i] += (int)w;
        ys[i] %= (int)h; if (ys[i] < 0) ys[i] += (int)h;
    }
    v3 c00 = v3load(rgb + 3 * ((size_t)ys[0] * w + xs[0]));
    v3 c10 = v3load(rgb + 3 * ((size_t)ys[0] * w + xs[1]));
    v3 c01 = v3load(rgb + 3 * ((size_t)ys[1] * w + xs[0]));
    v3 c11 = v3load(rgb + 3 * ((size_t)ys[1] * w + xs[1]));
    return lerp3(lerp3(c00, c10, fx), lerp3(c01, c11, fx), fy);
}

/* computeEnvironmentUV, Environment.slang:62-81 */
static void env_uv(const float* xform, int use_screen, v3 dir, uint32_t px, uint32_t py,
                   uint32_t W, uint32_t H, float uv[2])
{
    if (use_screen) {
        uv[0] = (float)px / (float)W;
        uv[1] = (float)py / (float)H;
        return;
    }
    v3 d = normalize(xform4_dir(xform, dir));
    float df[3] = { d.x, d.y, d.z };
    ora_dir_to_latlong(df, uv);
}

/* evaluateEnvironment, Environment.slang:98-133 */
void ora_env_evaluate(const ora_env* env, const float dirf[3], int as_background,
                      uint32_t px, uint32_t py, uint32_t W, uint32_t H, float out[3])
{
    v3 dir = v3load(dirf);
    float uv[2];
    v3 r;
    if (as_background) {
        env_uv(env->bg_xform, env->bg_use_screen, dir, px, py, W, H, uv);
        if (env->bg_rgb)
            r = tex_lookup(env->bg_rgb, env->bg_w, env->bg_h, uv[0], uv[1], env->filter);
        else
            r = lerp3(v3load(env->bg_top), v3load(env->bg_bottom), uv[1]);
    } else {
        env_uv(env->light_xform, 0, dir, px, py, W, H, uv);
        if (env->light_rgb)
            r = tex_lookup(env->light_rgb, env->light_w, env->light_h, uv[0], uv[1], env->filter);
        else
            r = lerp3(v3load(env->light_top), v3load(env->light_bottom), uv[1]);
    }
    v3store(out, r);
}

/* sampleEnvironment, Environment.slang:137-186 */
void ora_env_sample(const ora_env* env, const float rnd[2], float Lout[3], float* pdf, float out[3])
{
    if (env->light_rgb) {
        uint32_t w = env->light_w, h = env->light_h, count = w * h;
        uint32_t index = (uint32_t)((float)count * rnd[0]);
        if (index > count - 1) index = count - 1;
        ora_alias_entry e = env->alias[index];
        if (e.prob < rnd[1]) {
            index = e.alias;
            e = env->alias[index];
        }
        *pdf = e.pdf;
        uint32_t xi = index % w, yi = index / w;
        float uv[2] = { ((float)xi + 0.5f) / (float)w, ((float)yi + 0.5f) / (float)h };
        float d[3];
        ora_latlong_to_dir(uv, d);
        v3 L = normalize(xform4_dir(env->light_xform_inv, v3load(d)));
        v3store(Lout, L);
        v3store(out, v3load(env->light_rgb + 3 * ((size_t)yi * w + xi)));
    } else {
        /* sampleUniformDirection, Sampling.slang:20-34 */
        float cosTheta = 1.0f - 2.0f * rnd[1];
        float sinTheta = sqrtf(1.0f - cosTheta * cosTheta);
        float phi = 2.0f * ORA_PI * rnd[0];
        v3 L = V3(sinTheta * cosf(phi), sinTheta * sinf(phi), cosTheta);
        *pdf = 0.25f * ORA_PI_INV;
        L = normalize(xform4_dir(env->light_xform_inv, L));
        float Lf[3] = { L.x, L.y, L.z }, uv[2];
        ora_dir_to_latlong(Lf, uv);
        v3store(Lout, L);
        v3store(out, lerp3(v3load(env->light_top), v3load(env->light_bottom), uv[1]));
    }
}

/* AliasMap::build, Source/AliasMap.cpp:31-166.  Same operation order: float
 * luminance, float solid angle per row with the latitude stepped down from
 * pi/2, float running integral, the small/large two-cursor pass. */
void ora_alias_build(const float* rgb, uint32_t w, uint32_t h, ora_alias_entry* out, float* integral_out)
{
    uint32_t n = w * h;
    float* lum = (float*)malloc(sizeof(float) * n);
    float* lumArea = (float*)malloc(sizeof(float) * n);
    uint32_t* indexMap = (uint32_t*)malloc(sizeof(uint32_t) * n);
    float integral = 0.0f;
    float lonIncrement = (float)(2.0f * 3.14159265358979323846 / w);
    float latIncrement = (float)(3.14159265358979323846 / h);
    float latAngle = (float)(3.14159265358979323846 / 2.0);
    size_t li = 0;
    const float* p = rgb;
    for (uint32_t y = 0; y < h; y++) {
        float solidAngle = (sinf(latAngle) - sinf(latAngle - latIncrement)) * lonIncrement;
        latAngle -= latIncrement;
        for (uint32_t x = 0; x < w; x++) {
            float l = p[0] * 0.2125f + p[1] * 0.7154f + p[2] * 0.0721f;
            float la = solidAngle * l;
            lum[li] = l;
            lumArea[li] = la;
            integral += la;
            p += 3;
            li++;
        }
    }
    float average = integral / (float)n;
    uint32_t sml = 0, lrg = n;
    for (uint32_t i = 0; i < n; i++) {
        out[i].prob = lumArea[i] / average;
        out[i].pdf = lum[i] / integral;
        out[i].pad = 0.0f;
        uint32_t index = out[i].prob < 1.0f ? sml++ : --lrg;
        indexMap[index] = i;
        out[i].alias = i;
    }
    for (sml = 0; sml < lrg && lrg < n; sml++) {
        uint32_t is = indexMap[sml], il = indexMap[lrg];
        out[is].alias = il;
        out[il].prob -= 1.0f - out[is].prob;
        lrg += out[il].prob < 1.0f ? 1 : 0;
    }
    *integral_out = integral;
    free(lum);
    free(lumArea);
    free(indexMap);
}

/* PostProcessing.glsl:37-64 (brightness, optional ACES, sRGB approximation, alpha = 1). */
void ora_post(const float* accum, uint32_t n, const float brightness[3], int tone_map, int gamma, float* out)
{
    for (uint32_t i = 0; i < n; i++) {
        v3 c = mul(v3load(accum + 4 * i), v3load(brightness));
        if (tone_map) {
            /* toneMapACES, Colors.slang:45-53 */
            v3 num = mul(c, add(scale(c, 2.51f), v3s(0.03f)));
            v3 den = add(mul(c, add(scale(c, 2.43f), v3s(0.59f))), v3s(0.14f));
            v3 r = divv(num, den);
            c = V3(saturatef(r.x), saturatef(r.y), saturatef(r.z));
        }
        if (gamma) {
            /* linearTosRGB, Colors.slang:33-42, applied to the clamped colour */
            c = V3(saturatef(c.x), saturatef(c.y), saturatef(c.z));
            v3 s1 = V3(sqrtf(c.x), sqrtf(c.y), sqrtf(c.z));
            v3 s2 = V3(sqrtf(s1.x), sqrtf(s1.y), sqrtf(s1.z));
            v3 s3 = V3(sqrtf(s2.x), sqrtf(s2.y), sqrtf(s2.z));
            c = sub(add(scale(s1, 0.662002687f), scale(s2, 0.684122060f)),
                    add(scale(s3, 0.323583601f), scale(c, 0.0225411470f)));
        }
        out[4 * i + 0] = c.x;
        out[4 * i + 1] = c.y;
        out[4 * i + 2] = c.z;
        out[4 * i + 3] = 1.0f;
    }
}
