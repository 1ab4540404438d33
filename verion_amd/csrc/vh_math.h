// vh_math.h - device vector type and single-precision elementary functions.
//
// The transcendental calls of shaders/include/pt_core.slang (cos/sin :118-119, :171-172,
// :301-305, :658-659; acos/atan2 :204-205, :252-253; pow :452; log2 :338) use the Cephes method
// (S. L. Moshier, "Methods and Programs for Mathematical Functions", 1989: sinf.c, asinf.c,
// atanf.c, log2f.c, exp2f.c): Cody-Waite reduction plus the published minimax polynomials, in
// float, with contraction off for the whole library. Division and sqrt are the IEEE-rounded
// operations. The results are therefore a fixed function of the inputs, which the CPU oracle
// reproduces bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define VH_HD __host__ __device__ __forceinline__

namespace vh {

constexpr float kPi = 3.14159265358979f;
constexpr float kPiO2 = 1.5707963267948966192f;
constexpr float kPiO4 = 0.7853981633974483096f;

VH_HD float asf(uint32_t u) { return __builtin_bit_cast(float, u); }
VH_HD uint32_t asu(float f) { return __builtin_bit_cast(uint32_t, f); }

struct V3 {
    float x, y, z;
};
VH_HD V3 v3(float x, float y, float z) { return V3{x, y, z}; }
VH_HD V3 v3(float s) { return V3{s, s, s}; }
VH_HD V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
VH_HD V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
VH_HD V3 operator*(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
VH_HD V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
VH_HD V3 operator*(float s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
VH_HD V3 operator/(V3 a, float s) { return V3{a.x / s, a.y / s, a.z / s}; }
VH_HD V3 operator+(V3 a, float s) { return V3{a.x + s, a.y + s, a.z + s}; }
VH_HD V3 operator-(V3 a, float s) { return V3{a.x - s, a.y - s, a.z - s}; }
VH_HD V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
VH_HD V3 operator-(float s, V3 a) { return V3{s - a.x, s - a.y, s - a.z}; }
VH_HD V3 operator+(float s, V3 a) { return V3{s + a.x, s + a.y, s + a.z}; }
VH_HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
VH_HD V3 cross(V3 a, V3 b) { return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
VH_HD float length(V3 a) { return sqrtf(dot(a, a)); }
VH_HD V3 normalize(V3 a)
{
    float l = sqrtf(dot(a, a));
    return V3{a.x / l, a.y / l, a.z / l};
}
VH_HD float saturate(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }
VH_HD float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
VH_HD float lerp(float a, float b, float t) { return a + t * (b - a); }
VH_HD V3 lerp(V3 a, V3 b, float t) { return V3{lerp(a.x, b.x, t), lerp(a.y, b.y, t), lerp(a.z, b.z, t)}; }
VH_HD bool bad(V3 a)
{
    return __builtin_isnan(a.x) || __builtin_isnan(a.y) || __builtin_isnan(a.z) || __builtin_isinf(a.x) ||
           __builtin_isinf(a.y) || __builtin_isinf(a.z);
}
// uint(x) as the shader converts: negative and NaN give 0, saturating at the top.
VH_HD uint32_t f2u(float x) { return x > 0.0f ? (x >= 4294967040.0f ? 0xffffffffu : (uint32_t)x) : 0u; }
VH_HD uint32_t umin(uint32_t a, uint32_t b) { return a < b ? a : b; }
VH_HD uint32_t umax(uint32_t a, uint32_t b) { return a > b ? a : b; }

// sin / cos with the octant reduction of Cephes sinf.c; |x| < 8192.
VH_HD float sincos_poly(float x, bool cosine)
{
    float sign = 1.0f;
    if (x < 0.0f) {
        x = -x;
        if (!cosine)
            sign = -1.0f;
    }
    int j = (int)(x * 1.27323954473516f);
    float y = (float)j;
    if (j & 1) {
        j += 1;
        y += 1.0f;
    }
    j &= 7;
    if (cosine) {
        if (j > 3) {
            sign = -sign;
            j -= 4;
        }
        if (j > 1)
            sign = -sign;
    } else if (j > 3) {
        sign = -sign;
        j -= 4;
    }
    x = ((x - y * 0.78515625f) - y * 2.4187564849853515625e-4f) - y * 3.77489497744594108e-8f;
    float z = x * x;
    bool useCos = cosine ? !(j == 1 || j == 2) : (j == 1 || j == 2);
    float r;
    if (useCos)
        r = ((2.443315711809948e-5f * z - 1.388731625493765e-3f) * z + 4.166664568298827e-2f) * z * z - 0.5f * z + 1.0f;
    else
        r = ((-1.9515295891e-4f * z + 8.3321608736e-3f) * z - 1.6666654611e-1f) * z * x + x;
    return sign * r;
}
VH_HD float vsin(float x) { return sincos_poly(x, false); }
VH_HD float vcos(float x) { return sincos_poly(x, true); }

VH_HD float vasin(float x)
{
    float sign = 1.0f;
    float a = x;
    if (a < 0.0f) {
        sign = -1.0f;
        a = -a;
    }
    if (a > 1.0f)
        return 0.0f;
    float z;
    bool flag = false;
    if (a > 0.5f) {
        z = 0.5f * (1.0f - a);
        x = sqrtf(z);
        flag = true;
    } else {
        x = a;
        z = x * x;
    }
    z = ((((4.2163199048e-2f * z + 2.4181311049e-2f) * z + 4.5470025998e-2f) * z + 7.4953002686e-2f) * z + 1.6666752422e-1f) * z * x + x;
    if (flag) {
        z = z + z;
        z = kPiO2 - z;
    }
    return sign * z;
}

VH_HD float vacos(float x)
{
    if (x < -0.5f)
        return kPi - 2.0f * vasin(sqrtf(0.5f * (1.0f + x)));
    if (x > 0.5f)
        return 2.0f * vasin(sqrtf(0.5f * (1.0f - x)));
    return kPiO2 - vasin(x);
}

VH_HD float vatan(float x)
{
    float sign = 1.0f;
    if (x < 0.0f) {
        sign = -1.0f;
        x = -x;
    }
    float y;
    if (x > 2.414213562373095f) {
        y = kPiO2;
        x = -(1.0f / x);
    } else if (x > 0.4142135623730950f) {
        y = kPiO4;
        x = (x - 1.0f) / (x + 1.0f);
    } else {
        y = 0.0f;
    }
    float z = x * x;
    y += (((8.05374449538e-2f * z - 1.38776856032e-1f) * z + 1.99777106478e-1f) * z - 3.33329491539e-1f) * z * x + x;
    return sign * y;
}

VH_HD float vatan2(float y, float x)
{
    int code = 0;
    if (x < 0.0f)
        code = 2;
    if (y < 0.0f)
        code |= 1;
    if (x == 0.0f) {
        if (code & 1)
            return -kPiO2;
        if (y == 0.0f)
            return 0.0f;
        return kPiO2;
    }
    if (y == 0.0f)
        return (code & 2) ? kPi : 0.0f;
    float w = code == 2 ? kPi : (code == 3 ? -kPi : 0.0f);
    return w + vatan(y / x);
}

VH_HD float vlog2(float x)
{
    if (!(x > 0.0f))
        return x == 0.0f ? -__builtin_inff() : __builtin_nanf("");
    uint32_t b = asu(x);
    int e;
    if ((b & 0x7f800000u) == 0) {
        x = x * 16777216.0f;
        b = asu(x);
        e = (int)((b >> 23) & 0xff) - 126 - 24;
    } else {
        e = (int)((b >> 23) & 0xff) - 126;
    }
    x = asf((b & 0x807fffffu) | 0x3f000000u);
    if (x < 0.70710678118654752440f) {
        e -= 1;
        x = x + x - 1.0f;
    } else {
        x = x - 1.0f;
    }
    float z = x * x;
    float y = ((((((((7.0376836292e-2f * x - 1.1514610310e-1f) * x + 1.1676998740e-1f) * x - 1.2420140846e-1f) * x +
                   1.4249322787e-1f) * x - 1.6668057665e-1f) * x + 2.0000714765e-1f) * x - 2.4999993993e-1f) * x +
               3.3333331174e-1f) * x * z;
    y += -0.5f * z;
    z = y * 0.44269504088896340735992f;
    z += x * 0.44269504088896340735992f;
    z += y;
    z += x;
    z += (float)e;
    return z;
}

VH_HD float vexp2(float x)
{
    if (x > 127.0f)
        return __builtin_inff();
    if (x < -126.0f)
        return 0.0f;
    float px = floorf(x);
    int i = (int)px;
    x = x - px;
    if (x > 0.5f) {
        i += 1;
        x -= 1.0f;
    }
    px = x * (((((1.535336188319500e-4f * x + 1.339887440266574e-3f) * x + 9.618437357674640e-3f) * x +
                5.550332471162809e-2f) * x + 2.402264791363012e-1f) * x + 6.931472028550421e-1f);
    px += 1.0f;
    int i1 = i / 2;
    int i2 = i - i1;
    px = px * asf((uint32_t)(i1 + 127) << 23);
    px = px * asf((uint32_t)(i2 + 127) << 23);
    return px;
}

VH_HD float vpow(float x, float y)
{
    if (x <= 0.0f)
        return y > 0.0f ? 0.0f : (y == 0.0f ? 1.0f : __builtin_inff());
    if (x == 1.0f)
        return 1.0f;
    return vexp2(y * vlog2(x));
}

// pcg (pt_core.slang:69-74)
VH_HD uint32_t pcg(uint32_t v)
{
    uint32_t s = v * 747796405u + 2891336453u;
    uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
    return (w >> 22u) ^ w;
}

// Rng::next (pt_core.slang:79-83)
VH_HD float rnext(uint32_t& st)
{
    st = pcg(st);
    return (float)(st >> 8) * (1.0f / 16777216.0f);
}

VH_HD float lum(V3 c) { return dot(c, v3(0.2126f, 0.7152f, 0.0722f)); }

} // namespace vh
