/* Small FP32 vector helpers for the oracle (test infrastructure only). */
#ifndef ORA_MATH_H
#define ORA_MATH_H

#include <math.h>

typedef struct { float x, y, z; } v3;
typedef struct { float x, y; } v2;

#define ORA_PI 3.1415926535897932384626433832795f
#define ORA_PI_INV 0.318309886183790671537767526745f
#define ORA_FLOAT_EPS 0.000001f
#define ORA_RAY_TMIN 0.01f

static inline v3 V3(float x, float y, float z) { v3 r = { x, y, z }; return r; }
static inline v3 v3s(float s) { return V3(s, s, s); }
static inline v3 v3load(const float* p) { return V3(p[0], p[1], p[2]); }
static inline void v3store(float* p, v3 a) { p[0] = a.x; p[1] = a.y; p[2] = a.z; }
static inline v3 add(v3 a, v3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 mul(v3 a, v3 b) { return V3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 divv(v3 a, v3 b) { return V3(a.x / b.x, a.y / b.y, a.z / b.z); }
static inline v3 scale(v3 a, float s) { return V3(a.x * s, a.y * s, a.z * s); }
static inline v3 divs(v3 a, float s) { return V3(a.x / s, a.y / s, a.z / s); }
static inline v3 neg(v3 a) { return V3(-a.x, -a.y, -a.z); }
static inline float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 cross(v3 a, v3 b)
{
    return V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline float length3(v3 a) { return sqrtf(dot(a, a)); }
static inline v3 normalize(v3 a)
{
    /* HLSL normalize: v * rsqrt(dot(v,v)); computed as v / sqrt(dot) here. */
    return divs(a, sqrtf(dot(a, a)));
}
static inline float lerpf(float a, float b, float t) { return a + (b - a) * t; }
static inline v3 lerp3(v3 a, v3 b, float t) { return add(a, scale(sub(b, a), t)); }
static inline v3 lerp3v(v3 a, v3 b, v3 t) { return add(a, mul(sub(b, a), t)); }
static inline float clampf(float x, float lo, float hi) { return x < lo ? lo : (x > hi ? hi : x); }
static inline float saturatef(float x) { return clampf(x, 0.0f, 1.0f); }
static inline float maxf(float a, float b) { return a > b ? a : b; }
static inline float minf(float a, float b) { return a < b ? a : b; }
static inline float sqrf(float x) { return x * x; }
static inline int is_black(v3 c) { return c.x == 0.0f && c.y == 0.0f && c.z == 0.0f; }
static inline float luminance(v3 c) { return c.x * 0.2125f + c.y * 0.7154f + c.z * 0.0721f; }
static inline v3 pow3(v3 a, float e) { return V3(powf(a.x, e), powf(a.y, e), powf(a.z, e)); }
static inline v3 reflect3(v3 i, v3 n) { return sub(i, scale(n, 2.0f * dot(n, i))); }
/* HLSL refract(i, n, eta) */
static inline v3 refract3(v3 i, v3 n, float eta)
{
    float cosi = dot(neg(i), n);
    float k = 1.0f - eta * eta * (1.0f - cosi * cosi);
    if (k < 0.0f) return v3s(0.0f);
    /* eta*i + (eta*cosi - sqrt(k))*n */
    return add(scale(i, eta), scale(n, eta * cosi - sqrtf(k)));
}
/* 3x4 row-major affine transforms */
static inline v3 xform_point(const float* m, v3 p)
{
    return V3(m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3],
              m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
              m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11]);
}
static inline v3 xform_vector(const float* m, v3 p)
{
    return V3(m[0] * p.x + m[1] * p.y + m[2] * p.z,
              m[4] * p.x + m[5] * p.y + m[6] * p.z,
              m[8] * p.x + m[9] * p.y + m[10] * p.z);
}
/* 4x4 row-major applied to (v, 0) */
static inline v3 xform4_dir(const float* m, v3 p)
{
    return V3(m[0] * p.x + m[1] * p.y + m[2] * p.z,
              m[4] * p.x + m[5] * p.y + m[6] * p.z,
              m[8] * p.x + m[9] * p.y + m[10] * p.z);
}

#endif
