// Small vector math used by the gfx950 kernels (host + device).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define APH_HD __host__ __device__ __forceinline__

namespace aph {

constexpr float kPi = 3.1415926535897932384626433832795f;
constexpr float kPiInv = 0.318309886183790671537767526745f;
constexpr float kFloatEps = 0.000001f;  // M_FLOAT_EPS, Globals.slang:176-178
constexpr float kRayTMin = 0.01f;       // M_RAY_TMIN, Globals.slang:179

struct f3 {
    float x, y, z;
};

APH_HD f3 mk(float x, float y, float z) { return f3{x, y, z}; }
APH_HD f3 splat(float s) { return f3{s, s, s}; }
APH_HD f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
APH_HD f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
APH_HD f3 operator*(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
APH_HD f3 operator/(f3 a, f3 b) { return f3{a.x / b.x, a.y / b.y, a.z / b.z}; }
APH_HD f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
APH_HD f3 operator*(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
APH_HD f3 operator/(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
APH_HD f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
APH_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
APH_HD f3 cross(f3 a, f3 b) { return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
APH_HD float length(f3 a) { return sqrtf(dot(a, a)); }
APH_HD f3 normalize(f3 a) { return a / sqrtf(dot(a, a)); }
APH_HD float lerp(float a, float b, float t) { return a + (b - a) * t; }
APH_HD f3 lerp(f3 a, f3 b, float t) { return a + (b - a) * t; }
APH_HD f3 lerp(f3 a, f3 b, f3 t) { return a + (b - a) * t; }
APH_HD float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
APH_HD float sat(float x) { return clampf(x, 0.0f, 1.0f); }
APH_HD float sqr(float x) { return x * x; }
APH_HD bool isBlack(f3 c) { return c.x == 0.0f && c.y == 0.0f && c.z == 0.0f; }
APH_HD float luminance(f3 c) { return c.x * 0.2125f + c.y * 0.7154f + c.z * 0.0721f; }  // Colors.slang:26-29
APH_HD f3 fabs3(f3 a) { return f3{fabsf(a.x), fabsf(a.y), fabsf(a.z)}; }
APH_HD f3 pow3(f3 a, float e) { return f3{powf(a.x, e), powf(a.y, e), powf(a.z, e)}; }
APH_HD f3 reflect(f3 i, f3 n) { return i - n * (2.0f * dot(n, i)); }
APH_HD f3 refract(f3 i, f3 n, float eta)
{
    float cosi = dot(-i, n);
    float k = 1.0f - eta * eta * (1.0f - cosi * cosi);
    if (k < 0.0f) return splat(0.0f);
    return i * eta + n * (eta * cosi - sqrtf(k));
}
APH_HD float maxc(f3 a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }

// Row-major 3x4 affine transform.
struct Xf34 {
    float m[12];
};
APH_HD f3 xfPoint(const float* m, f3 p)
{
    return f3{m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3], m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
              m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11]};
}
APH_HD f3 xfVector(const float* m, f3 p)
{
    return f3{m[0] * p.x + m[1] * p.y + m[2] * p.z, m[4] * p.x + m[5] * p.y + m[6] * p.z,
              m[8] * p.x + m[9] * p.y + m[10] * p.z};
}

}  // namespace aph
