/* Oracle: temporal resolve, motion vectors and camera jitter (test infrastructure only).
 *
 * Restates DirectX/Shaders/TemporalResolve.hlsl:156-249 (with its helpers at
 * :90-154 and Colors.hlsli:42-56), the motion-vector helpers of
 * Shaders/GlobalPipelineState.slang:126-172 and the Halton jitter of
 * DirectX/PTRenderer.cpp:81-92,159-174, in plain C with no FMA contraction. */
#include "ora.h"

#include <math.h>
#include <stddef.h>

typedef struct { float x, y, z; } f3c;

static f3c ycocg(f3c c) /* RGBtoYCoCg, Colors.hlsli:43-49 */
{
    f3c r = { (0.25f * c.x + 0.5f * c.y) + 0.25f * c.z, 0.5f * c.x - 0.5f * c.z,
              (-0.25f * c.x + 0.5f * c.y) - 0.25f * c.z };
    return r;
}

static f3c ycocg_inv(f3c c) /* YCoCgtoRGB, Colors.hlsli:52-56 */
{
    float t = c.x - c.z;
    f3c r = { t + c.y, c.x + c.z, t - c.y };
    return r;
}

static float clampf01(float v) { return v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v); }

/* sampleHistoryCatmullRom, TemporalResolve.hlsl:97-131 */
static f3c catmull_rom(const float* hist, float px, float py, int W, int H)
{
    float sx = px - 0.5f, sy = py - 0.5f;
    float bx = floorf(sx), by = floorf(sy);
    float f[2] = { sx - bx, sy - by };
    float w[2][4];
    for (int a = 0; a < 2; a++) {
        float t = f[a], t2 = t * t, t3 = t2 * t;
        w[a][0] = (-0.5f * t3 + t2) - 0.5f * t;
        w[a][1] = (1.5f * t3 - 2.5f * t2) + 1.0f;
        w[a][2] = (-1.5f * t3 + 2.0f * t2) + 0.5f * t;
        w[a][3] = 0.5f * t3 - 0.5f * t2;
    }
    f3c r = { 0.0f, 0.0f, 0.0f };
    /* taps clamp to the image: a base beyond [-4, size + 4] reads the same texels */
    int ix = (int)fminf(fmaxf(bx, -4.0f), (float)(W + 4)), iy = (int)fminf(fmaxf(by, -4.0f), (float)(H + 4));
    for (int j = 0; j < 4; j++) {
        for (int i = 0; i < 4; i++) {
            int tx = ix + i - 1, ty = iy + j - 1;
            tx = tx < 0 ? 0 : (tx > W - 1 ? W - 1 : tx);
            ty = ty < 0 ? 0 : (ty > H - 1 ? H - 1 : ty);
            const float* h = hist + 4 * ((size_t)ty * W + tx);
            float ww = w[0][i] * w[1][j];
            r.x += h[0] * ww;
            r.y += h[1] * ww;
            r.z += h[2] * ww;
        }
    }
    r.x = r.x > 0.0f ? r.x : 0.0f;
    r.y = r.y > 0.0f ? r.y : 0.0f;
    r.z = r.z > 0.0f ? r.z : 0.0f;
    return r;
}

/* clipToBox, TemporalResolve.hlsl:138-152 */
static f3c clip_to_box(f3c h, f3c c, f3c lo, f3c hi)
{
    float d[3] = { h.x - c.x, h.y - c.y, h.z - c.z }, cc[3] = { c.x, c.y, c.z };
    float l[3] = { lo.x, lo.y, lo.z }, u[3] = { hi.x, hi.y, hi.z };
    float tm = INFINITY;
    for (int k = 0; k < 3; k++) {
        float mag = fabsf(d[k]);
        float lim = d[k] > 0.0f ? u[k] - cc[k] : cc[k] - l[k];
        float t = mag > 1e-6f ? lim / (mag > 1e-6f ? mag : 1e-6f) : 1.0f;
        if (t < tm) tm = t;
    }
    float s = clampf01(tm);
    f3c r = { c.x + d[0] * s, c.y + d[1] * s, c.z + d[2] * s };
    return r;
}

/* TemporalResolve, TemporalResolve.hlsl:156-249, for one pixel */
static void resolve_pixel(const float* accum, const float* depth, const float* motion, const float* hist, int W, int H,
                          int isReset, float amin, float amax, float speedRef, int x, int y, float out[4])
{
    size_t c = (size_t)y * W + x;
    const float* cur = accum + 4 * c;
    f3c curY = ycocg((f3c){ cur[0], cur[1], cur[2] });
    f3c m1 = { 0, 0, 0 }, m2 = { 0, 0, 0 }, nmin = { 1e30f, 1e30f, 1e30f }, nmax = { -1e30f, -1e30f, -1e30f };
    for (int j = -1; j <= 1; j++) {
        for (int i = -1; i <= 1; i++) {
            f3c v;
            if (i == 0 && j == 0) {
                v = curY;
            } else {
                int nx = x + i, ny = y + j;
                nx = nx < 0 ? 0 : (nx > W - 1 ? W - 1 : nx);
                ny = ny < 0 ? 0 : (ny > H - 1 ? H - 1 : ny);
                const float* a = accum + 4 * ((size_t)ny * W + nx);
                v = ycocg((f3c){ a[0], a[1], a[2] });
            }
            m1.x += v.x; m1.y += v.y; m1.z += v.z;
            m2.x += v.x * v.x; m2.y += v.y * v.y; m2.z += v.z * v.z;
            nmin.x = fminf(nmin.x, v.x); nmin.y = fminf(nmin.y, v.y); nmin.z = fminf(nmin.z, v.z);
            nmax.x = fmaxf(nmax.x, v.x); nmax.y = fmaxf(nmax.y, v.y); nmax.z = fmaxf(nmax.z, v.z);
        }
    }
    f3c mean = { m1.x / 9.0f, m1.y / 9.0f, m1.z / 9.0f };
    f3c sg = { sqrtf(fmaxf(m2.x / 9.0f - mean.x * mean.x, 0.0f)), sqrtf(fmaxf(m2.y / 9.0f - mean.y * mean.y, 0.0f)),
               sqrtf(fmaxf(m2.z / 9.0f - mean.z * mean.z, 0.0f)) };
    f3c lo = { fmaxf(mean.x - sg.x, nmin.x), fmaxf(mean.y - sg.y, nmin.y), fmaxf(mean.z - sg.z, nmin.z) };
    f3c hi = { fminf(mean.x + sg.x, nmax.x), fminf(mean.y + sg.y, nmax.y), fminf(mean.z + sg.z, nmax.z) };
    lo.x = fminf(lo.x, curY.x); lo.y = fminf(lo.y, curY.y); lo.z = fminf(lo.z, curY.z);
    hi.x = fmaxf(hi.x, curY.x); hi.y = fmaxf(hi.y, curY.y); hi.z = fmaxf(hi.z, curY.z);
    const float* mv = motion + 4 * c;
    float px = ((float)x + 0.5f) + mv[0], py = ((float)y + 0.5f) + mv[1];
    int valid = isReset == 0;
    valid = valid && px >= 0.0f && py >= 0.0f && px < (float)W && py < (float)H;
    float vz = fabsf(depth[c]);
    valid = valid && (fabsf(mv[2]) < 0.05f * fmaxf(vz, 1e-4f));
    if (!(isfinite(px) && isfinite(py))) px = py = 0.0f; /* already invalid */
    f3c h = catmull_rom(hist, px, py, W, H);
    h = ycocg_inv(clip_to_box(ycocg(h), curY, lo, hi));
    valid = valid && isfinite(h.x) && isfinite(h.y) && isfinite(h.z);
    float speed = sqrtf(mv[0] * mv[0] + mv[1] * mv[1]);
    float s = clampf01(speed / fmaxf(speedRef, 1e-4f));
    float alpha = amin + (amax - amin) * s;
    alpha = valid ? clampf01(alpha) : 1.0f;
    out[0] = h.x + (cur[0] - h.x) * alpha;
    out[1] = h.y + (cur[1] - h.y) * alpha;
    out[2] = h.z + (cur[2] - h.z) * alpha;
    out[3] = cur[3];
}

void ora_temporal_resolve(const float* accum, const float* depth, const float* motion, const float* hist_in,
                          float* hist_out, uint32_t W, uint32_t H, int is_reset, float alpha_min, float alpha_max,
                          float speed_reference)
{
    for (uint32_t y = 0; y < H; y++)
        for (uint32_t x = 0; x < W; x++)
            resolve_pixel(accum, depth, motion, hist_in, (int)W, (int)H, is_reset, alpha_min, alpha_max,
                          speed_reference, (int)x, (int)y, hist_out + 4 * ((size_t)y * W + x));
}

/* computeViewZ, GlobalPipelineState.slang:142-149 */
float ora_view_z(const float inv_view[16], const float p[3])
{
    float f[3] = { -inv_view[8], -inv_view[9], -inv_view[10] };
    float d[3] = { p[0] - inv_view[12], p[1] - inv_view[13], p[2] - inv_view[14] };
    return -((d[0] * f[0] + d[1] * f[1]) + d[2] * f[2]);
}

static void mul4(const float* M, const float v[4], float o[4])
{
    for (int i = 0; i < 4; i++)
        o[i] = ((M[4 * i] * v[0] + M[4 * i + 1] * v[1]) + M[4 * i + 2] * v[2]) + M[4 * i + 3] * v[3];
}

/* computeScreenMotion, GlobalPipelineState.slang:126-140: previous minus current pixel position */
void ora_screen_motion(const float view_proj[16], const float view_proj_prev[16], const float p[4], uint32_t W,
                       uint32_t H, float out[2])
{
    float c[4], q[4];
    mul4(view_proj, p, c);
    mul4(view_proj_prev, p, q);
    float cw = fmaxf(fabsf(c[3]), 1e-6f), qw = fmaxf(fabsf(q[3]), 1e-6f);
    float cx = c[0] / cw, cy = c[1] / cw, qx = q[0] / qw, qy = q[1] / qw;
    float csx = (cx * 0.5f + 0.5f) * (float)W, csy = (-cy * 0.5f + 0.5f) * (float)H;
    float qsx = (qx * 0.5f + 0.5f) * (float)W, qsy = (-qy * 0.5f + 0.5f) * (float)H;
    out[0] = qsx - csx;
    out[1] = qsy - csy;
}

/* getTemporalJitter / halton, PTRenderer.cpp:81-92,169-174 */
static float halton(uint32_t index, uint32_t base)
{
    float result = 0.0f, scale = 1.0f / (float)base;
    while (index > 0) {
        result += scale * (float)(index % base);
        index /= base;
        scale /= (float)base;
    }
    return result;
}

void ora_halton_jitter(uint32_t frame_index, uint32_t phase_count, float out[2])
{
    uint32_t index = (phase_count > 0u ? frame_index % phase_count : frame_index) + 1u;
    out[0] = halton(index, 2) - 0.5f;
    out[1] = halton(index, 3) - 0.5f;
}
