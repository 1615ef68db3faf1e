/*
 * luxion_detmath.h -- deterministic single-precision math for the luxion
 * MI355X path tracer.
 *
 * Every transcendental the shading path needs (sin, cos, atan, atan2, acos,
 * log2, exp2, pow) is defined here as a fixed sequence of IEEE-754 binary32
 * additions, multiplications, divisions, square roots and exact bit
 * operations.  Compiled with contraction off (-ffp-contract=off), correctly
 * rounded fp32 divide/sqrt and no fast-math, the sequence yields the same bits
 * on the gfx950 VALU (hipcc) and on an x86-64 host (g++ or hipcc host pass).
 * That is what lets the GPU renderer and the CPU restatement agree bit for
 * bit on hit records, survivor sets and per-pixel radiance.
 *
 * The polynomials and range reductions follow Stephen L. Moshier's Cephes
 * single-precision library (sinf.c, cosf.c, atanf.c, atan2f.c, asinf.c,
 * log2f.c, exp2f.c; "Cephes Math Library Release 2.2", 1984-1992), evaluated
 * with explicit Horner steps.  Accuracy against a double-precision libm is a
 * few ulp over the ranges the renderer uses; the unit tests state the bounds.
 *
 * The reference renderer (src/shaders/cook_torrance.cu, src/interactions.cu,
 * src/kernels/shade_path.cu, src/kernels/sample_direct_light.cu,
 * src/material_queries.cu) calls the CUDA device libm for these functions;
 * that library is closed and is not imitated here.
 */
#ifndef LUXION_DETMATH_H
#define LUXION_DETMATH_H

#include <stdint.h>

#if defined(__HIPCC__)
#define LXM_FN __host__ __device__ static inline
#else
#define LXM_FN static inline
#endif

#define LXM_PIF 3.14159265358979323846f
#define LXM_PIO2F 1.57079632679489661923f
#define LXM_PIO4F 0.78539816339744830962f

LXM_FN uint32_t lxm_f2u(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    return u;
}

LXM_FN float lxm_u2f(uint32_t u) {
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

LXM_FN float lxm_sqrt(float x) { return __builtin_sqrtf(x); }
LXM_FN float lxm_abs(float x) { return lxm_u2f(lxm_f2u(x) & 0x7fffffffu); }
LXM_FN float lxm_floor(float x) { return __builtin_floorf(x); }
LXM_FN int lxm_isnan(float x) { return x != x; }
LXM_FN int lxm_isfinite(float x) { return (lxm_f2u(x) & 0x7f800000u) != 0x7f800000u; }
/* glm::min / glm::max / glm::clamp semantics: x < y ? x : y, x > y ? x : y */
LXM_FN float lxm_min(float x, float y) { return x < y ? x : y; }
LXM_FN float lxm_max(float x, float y) { return x > y ? x : y; }
LXM_FN float lxm_clamp(float x, float lo, float hi) { return lxm_min(lxm_max(x, lo), hi); }

/* x * 2^e for e in [-252, 254], exact unless the result is subnormal. */
LXM_FN float lxm_ldexp(float x, int e) {
    if (e > 127) {
        x = x * lxm_u2f(0x7f000000u); /* 2^127 */
        e -= 127;
        if (e > 127) e = 127;
    } else if (e < -126) {
        x = x * lxm_u2f(0x00800000u); /* 2^-126 */
        e += 126;
        if (e < -126) e = -126;
    }
    return x * lxm_u2f((uint32_t)(e + 127) << 23);
}

/* Cephes sinf.c / cosf.c: reduction by pi/4 in three parts, then the
 * minimax sine or cosine polynomial on [-pi/4, pi/4]. */
#define LXM_DP1 0.78515625f
#define LXM_DP2 2.4187564849853515625e-4f
#define LXM_DP3 3.77489497744594108e-8f
#define LXM_FOPI 1.27323954473516f

LXM_FN float lxm_sinpoly_(float x, float z) {
    float y = -1.9515295891e-4f * z;
    y = y + 8.3321608736e-3f;
    y = y * z;
    y = y - 1.6666654611e-1f;
    y = y * z;
    y = y * x;
    return y + x;
}

LXM_FN float lxm_cospoly_(float z) {
    float y = 2.443315711809948e-5f * z;
    y = y - 1.388731625493765e-3f;
    y = y * z;
    y = y + 4.166664568298827e-2f;
    y = y * z;
    y = y * z;
    y = y - 0.5f * z;
    return y + 1.0f;
}

LXM_FN float lxm_sin(float xx) {
    if (lxm_isnan(xx)) return xx;
    float x = xx;
    int neg = 0;
    if (x < 0.0f) { neg = 1; x = -x; }
    if (x > 8388608.0f) return neg ? -0.0f : 0.0f; /* no precision left */
    int j = (int)(LXM_FOPI * x);
    float y = (float)j;
    if (j & 1) { j += 1; y += 1.0f; }
    j &= 7;
    if (j > 3) { neg = !neg; j -= 4; }
    x = ((x - y * LXM_DP1) - y * LXM_DP2) - y * LXM_DP3;
    float z = x * x;
    float r = (j == 1 || j == 2) ? lxm_cospoly_(z) : lxm_sinpoly_(x, z);
    return neg ? -r : r;
}

LXM_FN float lxm_cos(float xx) {
    if (lxm_isnan(xx)) return xx;
    float x = lxm_abs(xx);
    if (x > 8388608.0f) return 1.0f;
    int j = (int)(LXM_FOPI * x);
    float y = (float)j;
    if (j & 1) { j += 1; y += 1.0f; }
    j &= 7;
    int neg = 0;
    if (j > 3) { j -= 4; neg = !neg; }
    if (j > 1) neg = !neg;
    x = ((x - y * LXM_DP1) - y * LXM_DP2) - y * LXM_DP3;
    float z = x * x;
    float r = (j == 1 || j == 2) ? lxm_sinpoly_(x, z) : lxm_cospoly_(z);
    return neg ? -r : r;
}

/* Cephes atanf.c */
LXM_FN float lxm_atan(float xx) {
    float x = xx;
    int neg = 0;
    if (x < 0.0f) { neg = 1; x = -x; }
    float y;
    if (x > 2.414213562373095f) {
        y = LXM_PIO2F;
        x = -(1.0f / x);
    } else if (x > 0.4142135623730950f) {
        y = LXM_PIO4F;
        x = (x - 1.0f) / (x + 1.0f);
    } else {
        y = 0.0f;
    }
    float z = x * x;
    float p = 8.05374449538e-2f * z;
    p = p - 1.38776856032e-1f;
    p = p * z;
    p = p + 1.99777106478e-1f;
    p = p * z;
    p = p - 3.33329491539e-1f;
    p = p * z;
    p = p * x;
    p = p + x;
    y = y + p;
    return neg ? -y : y;
}

/* Cephes atan2f.c quadrant logic. */
LXM_FN float lxm_atan2(float y, float x) {
    if (lxm_isnan(x) || lxm_isnan(y)) return x + y;
    if (x == 0.0f) {
        if (y > 0.0f) return LXM_PIO2F;
        if (y < 0.0f) return -LXM_PIO2F;
        return 0.0f;
    }
    if (y == 0.0f) {
        return x > 0.0f ? 0.0f : LXM_PIF;
    }
    float w;
    if (x < 0.0f) w = (y < 0.0f) ? -LXM_PIF : LXM_PIF;
    else w = 0.0f;
    return w + lxm_atan(y / x);
}

/* Cephes asinf.c / acosf.c */
LXM_FN float lxm_asin(float xx) {
    float a = lxm_abs(xx);
    if (!(a <= 1.0f)) return lxm_u2f(0x7fc00000u);
    if (a < 1.0e-4f) return xx;
    float x, z;
    int flag;
    if (a > 0.5f) {
        z = 0.5f * (1.0f - a);
        x = lxm_sqrt(z);
        flag = 1;
    } else {
        x = a;
        z = x * x;
        flag = 0;
    }
    float p = 4.2163199048e-2f * z;
    p = p + 2.4181311049e-2f;
    p = p * z;
    p = p + 4.5470025998e-2f;
    p = p * z;
    p = p + 7.4953002686e-2f;
    p = p * z;
    p = p + 1.6666752422e-1f;
    p = p * z;
    p = p * x;
    z = p + x;
    if (flag) {
        z = z + z;
        z = LXM_PIO2F - z;
    }
    return xx < 0.0f ? -z : z;
}

LXM_FN float lxm_acos(float x) {
    if (!(x >= -1.0f && x <= 1.0f)) return lxm_u2f(0x7fc00000u);
    if (x < -0.5f) return LXM_PIF - 2.0f * lxm_asin(lxm_sqrt(0.5f * (1.0f + x)));
    if (x > 0.5f) return 2.0f * lxm_asin(lxm_sqrt(0.5f * (1.0f - x)));
    return LXM_PIO2F - lxm_asin(x);
}

/* Cephes log2f.c; x <= 0 gives -inf (x == 0) or NaN. */
LXM_FN float lxm_log2(float xx) {
    uint32_t u = lxm_f2u(xx);
    if (lxm_isnan(xx)) return xx;
    if (xx == 0.0f) return lxm_u2f(0xff800000u);
    if (xx < 0.0f) return lxm_u2f(0x7fc00000u);
    if (u == 0x7f800000u) return xx;
    int e;
    if ((u & 0x7f800000u) == 0u) { /* subnormal: scale up by 2^23 */
        float s = xx * 8388608.0f;
        u = lxm_f2u(s);
        e = (int)((u >> 23) & 0xffu) - 126 - 23;
    } else {
        e = (int)((u >> 23) & 0xffu) - 126;
    }
    float x = lxm_u2f((u & 0x807fffffu) | 0x3f000000u); /* mantissa in [0.5, 1) */
    if (x < 0.707106781186547524f) {
        e -= 1;
        x = x + x - 1.0f;
    } else {
        x = x - 1.0f;
    }
    float z = x * x;
    float y = 7.0376836292e-2f * x;
    y = y - 1.1514610310e-1f;
    y = y * x;
    y = y + 1.1676998740e-1f;
    y = y * x;
    y = y - 1.2420140846e-1f;
    y = y * x;
    y = y + 1.4249322787e-1f;
    y = y * x;
    y = y - 1.6668057665e-1f;
    y = y * x;
    y = y + 2.0000714765e-1f;
    y = y * x;
    y = y - 2.4999993993e-1f;
    y = y * x;
    y = y + 3.3333331174e-1f;
    y = y * x;
    y = y * z;
    y = y + (-0.5f * z);
    const float LOG2EA = 0.44269504088896340735992f;
    float r = y * LOG2EA;
    r = r + x * LOG2EA;
    r = r + y;
    r = r + x;
    r = r + (float)e;
    return r;
}

/* Cephes exp2f.c */
LXM_FN float lxm_exp2(float xx) {
    if (lxm_isnan(xx)) return xx;
    if (xx > 128.0f) return lxm_u2f(0x7f800000u);
    if (xx < -150.0f) return 0.0f;
    float px = lxm_floor(xx);
    int i = (int)px;
    float x = xx - px;
    if (x > 0.5f) {
        i += 1;
        x = x - 1.0f;
    }
    float p = 1.535336188319500e-4f * x;
    p = p + 1.339887440266574e-3f;
    p = p * x;
    p = p + 9.618437357674640e-3f;
    p = p * x;
    p = p + 5.550332471162809e-2f;
    p = p * x;
    p = p + 2.402264791363012e-1f;
    p = p * x;
    p = p + 6.931472028550421e-1f;
    p = p * x;
    p = 1.0f + p;
    return lxm_ldexp(p, i);
}

/* pow for the non-negative bases the renderer raises (texel gamma 2.2,
 * display gamma 0.45): pow(0, y>0) = 0, pow(1, y) = 1. */
LXM_FN float lxm_pow(float x, float y) {
    if (x == 1.0f || y == 0.0f) return 1.0f;
    if (x == 0.0f) return y > 0.0f ? 0.0f : lxm_u2f(0x7f800000u);
    if (x < 0.0f) return lxm_u2f(0x7fc00000u);
    return lxm_exp2(y * lxm_log2(x));
}

#endif /* LUXION_DETMATH_H */
