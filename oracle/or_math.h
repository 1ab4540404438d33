/*
 * or_math.h - ORACLE (test infrastructure only; never linked into the product library).
 *
 * Single-precision elementary functions for the CPU oracle, written from Stephen L. Moshier's
 * Cephes Math Library (sinf.c, asinf.c, atanf.c, log2f.c, exp2f.c; "Methods and Programs for
 * Mathematical Functions", 1989): Cody-Waite reduction plus the published minimax polynomials,
 * evaluated in float with no contraction. The GPU library implements the same published method, so
 * both sides produce identical bits; the transcendental calls of pt_core.slang (cos/sin at :118-119,
 * :171-172, :301-305, :658-659; acos/atan2 at :204-205, :252-253; pow at :452; log2 at :338) go
 * through these functions on both sides.
 *
 * Pinned by tests/test_oracle_units.py (test_sin_cos_against_float64, test_acos_atan2_against_float64,
 * test_log2_exp2_pow) against numpy's float64 functions (error bounds in ulps) and against closed
 * values (sin(pi/6), acos(0), log2(2^k), exp2(k)).
 */
#ifndef OR_MATH_H
#define OR_MATH_H

#include <stdint.h>
#include <string.h>

static inline uint32_t or_f2b(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float or_b2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

static const float OR_PI = 3.14159265358979f;
static const float OR_PIO2 = 1.5707963267948966192f;
static const float OR_PIO4 = 0.7853981633974483096f;

/* sin/cos with the Cephes octant reduction (sinf.c). Valid for |x| < 8192. */
static inline float or_sincos_poly(float x, int cosine)
{
    float sign = 1.0f;
    if (x < 0.0f) {
        x = -x;
        if (!cosine)
            sign = -1.0f;
    }
    int j = (int)(x * 1.27323954473516f); /* 4/pi */
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
    /* Extended-precision modular arithmetic: pi/4 = DP1 + DP2 + DP3. */
    x = ((x - y * 0.78515625f) - y * 2.4187564849853515625e-4f) - y * 3.77489497744594108e-8f;
    float z = x * x;
    float r;
    int use_cos = cosine ? !(j == 1 || j == 2) : (j == 1 || j == 2);
    if (use_cos)
        r = ((2.443315711809948e-5f * z - 1.388731625493765e-3f) * z + 4.166664568298827e-2f) * z * z - 0.5f * z + 1.0f;
    else
        r = ((-1.9515295891e-4f * z + 8.3321608736e-3f) * z - 1.6666654611e-1f) * z * x + x;
    return sign * r;
}
static inline float or_sinf(float x) { return or_sincos_poly(x, 0); }
static inline float or_cosf(float x) { return or_sincos_poly(x, 1); }

/* asinf.c */
static inline float or_asinf(float x)
{
    float sign = 1.0f;
    float a = x;
    if (a < 0.0f) {
        sign = -1.0f;
        a = -a;
    }
    if (a > 1.0f)
        return 0.0f; /* domain: callers clamp */
    float z;
    int flag = 0;
    if (a > 0.5f) {
        z = 0.5f * (1.0f - a);
        x = sqrtf(z);
        flag = 1;
    } else {
        x = a;
        z = x * x;
    }
    z = ((((4.2163199048e-2f * z + 2.4181311049e-2f) * z + 4.5470025998e-2f) * z + 7.4953002686e-2f) * z + 1.6666752422e-1f) * z * x + x;
    if (flag) {
        z = z + z;
        z = OR_PIO2 - z;
    }
    return sign * z;
}

static inline float or_acosf(float x)
{
    if (x < -0.5f)
        return OR_PI - 2.0f * or_asinf(sqrtf(0.5f * (1.0f + x)));
    if (x > 0.5f)
        return 2.0f * or_asinf(sqrtf(0.5f * (1.0f - x)));
    return OR_PIO2 - or_asinf(x);
}

/* atanf.c */
static inline float or_atanf(float x)
{
    float sign = 1.0f;
    if (x < 0.0f) {
        sign = -1.0f;
        x = -x;
    }
    float y;
    if (x > 2.414213562373095f) {
        y = OR_PIO2;
        x = -(1.0f / x);
    } else if (x > 0.4142135623730950f) {
        y = OR_PIO4;
        x = (x - 1.0f) / (x + 1.0f);
    } else {
        y = 0.0f;
    }
    float z = x * x;
    y += (((8.05374449538e-2f * z - 1.38776856032e-1f) * z + 1.99777106478e-1f) * z - 3.33329491539e-1f) * z * x + x;
    return sign * y;
}

/* atan2f.c */
static inline float or_atan2f(float y, float x)
{
    int code = 0;
    if (x < 0.0f)
        code = 2;
    if (y < 0.0f)
        code |= 1;
    if (x == 0.0f) {
        if (code & 1)
            return -OR_PIO2;
        if (y == 0.0f)
            return 0.0f;
        return OR_PIO2;
    }
    if (y == 0.0f) {
        if (code & 2)
            return OR_PI;
        return 0.0f;
    }
    float w;
    switch (code) {
    case 2: w = OR_PI; break;
    case 3: w = -OR_PI; break;
    default: w = 0.0f; break;
    }
    return w + or_atanf(y / x);
}

/* log2f.c, for finite x > 0 (0 gives -inf). */
static inline float or_log2f(float x)
{
    if (!(x > 0.0f))
        return x == 0.0f ? -1.0f / 0.0f : 0.0f / 0.0f;
    uint32_t b = or_f2b(x);
    int e;
    if ((b & 0x7f800000u) == 0) { /* subnormal: scale by 2^24 first */
        x = x * 16777216.0f;
        b = or_f2b(x);
        e = (int)((b >> 23) & 0xff) - 126 - 24;
    } else {
        e = (int)((b >> 23) & 0xff) - 126;
    }
    x = or_b2f((b & 0x807fffffu) | 0x3f000000u); /* mantissa in [0.5, 1) */
    if (x < 0.70710678118654752440f) {
        e -= 1;
        x = x + x - 1.0f;
    } else {
        x = x - 1.0f;
    }
    float z = x * x;
    float y = ((((((((7.0376836292e-2f * x - 1.1514610310e-1f) * x + 1.1676998740e-1f) * x - 1.2420140846e-1f) * x
                  + 1.4249322787e-1f) * x - 1.6668057665e-1f) * x + 2.0000714765e-1f) * x - 2.4999993993e-1f) * x
              + 3.3333331174e-1f) * x * z;
    y += -0.5f * z;
    /* log2(1+x) = (y + x) * log2(e), with log2(e) = 1 + LOG2EA */
    z = y * 0.44269504088896340735992f;
    z += x * 0.44269504088896340735992f;
    z += y;
    z += x;
    z += (float)e;
    return z;
}

/* exp2f.c */
static inline float or_exp2f(float x)
{
    if (x > 127.0f)
        return 1.0f / 0.0f;
    if (x < -126.0f)
        return 0.0f;
    float px = floorf(x);
    int i = (int)px;
    x = x - px;
    if (x > 0.5f) {
        i += 1;
        x -= 1.0f;
    }
    px = x * (((((1.535336188319500e-4f * x + 1.339887440266574e-3f) * x + 9.618437357674640e-3f) * x
               + 5.550332471162809e-2f) * x + 2.402264791363012e-1f) * x + 6.931472028550421e-1f);
    px += 1.0f;
    /* ldexp by building 2^i (i in [-126, 128]) in two steps so it stays normal */
    int i1 = i / 2;
    int i2 = i - i1;
    px = px * or_b2f((uint32_t)(i1 + 127) << 23);
    px = px * or_b2f((uint32_t)(i2 + 127) << 23);
    return px;
}

/* pow for x >= 0 (the only use: the sheen distribution, pt_core.slang:452). */
static inline float or_powf(float x, float y)
{
    if (x <= 0.0f)
        return y > 0.0f ? 0.0f : (y == 0.0f ? 1.0f : 1.0f / 0.0f);
    if (x == 1.0f)
        return 1.0f;
    return or_exp2f(y * or_log2f(x));
}

#endif
