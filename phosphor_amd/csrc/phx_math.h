/*
 * phx_math.h — single-precision elementary functions with fully specified
 * results, compiled identically for gfx950 device code and for x86-64 host
 * code.
 *
 * The reference evaluates cos/sin/exp/log/pow/acos/atan2/frexp/ldexp with the
 * OpenCL implementation's builtins (random.h:45,56; camera.h:39-40;
 * bsdfs.h:51-52,88-97; light_sampling.h:149-157,193,258; shading.h:128-132;
 * envmap.h:10-11; encoding.h:16-17,36; photon_hash.h:96 native_exp), whose
 * bits are implementation-defined.  Here each function is built only from
 * IEEE-754 add/sub/mul/div/sqrt, comparisons and bit manipulation, so with FP
 * contraction disabled (-ffp-contract=off on both compilers) and IEEE
 * denormals, the host and the GPU produce the same bits for the same input.
 *
 * Methods (all textbook; coefficients are Taylor-series coefficients written
 * as decimal fractions, reduction constants are the leading bits of π/2 and
 * ln 2 split into exactly representable pieces):
 *   sin/cos  Cody & Waite (1980) additive range reduction by π/2 with a
 *            three-piece split (11+11+24 bits), degree-9/10 Taylor kernels on
 *            [-π/4, π/4].  Valid for |x| < 8192·π/2 ≈ 12868.
 *   exp      reduction x = k·ln2 + r (two-piece ln2), degree-7 Taylor kernel,
 *            exact scaling by 2^k incl. gradual underflow.
 *   log      x = m·2^e with m ∈ [√½, √2), log m = 2·atanh(f/(2+f)) series.
 *   atan     reflection |x|>1 → π/2 − atan(1/x), then the tan(π/12)
 *            reduction atan x = π/6 + atan((√3x−1)/(x+√3)), Taylor kernel.
 *   atan2    quadrant logic over atan of the smaller ratio.
 *   acos     2·atan(√((1−x)/(1+x))).
 *   pow      exp(y·log x) for x ≥ 0 (used only by sRGB decoding).
 * Accuracy against correctly rounded results is checked by the CPU test
 * suite (≤ 2 ulp for sin/cos/exp/log in their working ranges).
 */
#ifndef PHX_MATH_H
#define PHX_MATH_H

#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#define PHX_HD static inline __host__ __device__
#else
#define PHX_HD static inline
#endif

#define PHX_PI 3.141592654f /* constants.h:14 */
#define PHX_EPS 1e-5f       /* constants.h:11 */
#define PHX_INF_F 1e9f      /* constants.h:13 (a large finite value, not IEEE inf) */

PHX_HD uint32_t phx_as_uint(float f) {
    union { float f; uint32_t u; } v;
    v.f = f;
    return v.u;
}

PHX_HD float phx_as_float(uint32_t u) {
    union { float f; uint32_t u; } v;
    v.u = u;
    return v.f;
}

PHX_HD float phx_sqrtf(float x) { return __builtin_sqrtf(x); }
PHX_HD float phx_fabsf(float x) { return __builtin_fabsf(x); }
PHX_HD float phx_floorf(float x) { return __builtin_floorf(x); }
PHX_HD float phx_ceilf(float x) { return __builtin_ceilf(x); }

/* fmax/fmin with one fixed answer for signed zeros (the second operand wins
 * on equality), so both compilers agree bit for bit; a NaN operand yields
 * the other operand, as C99 fmaxf does. */
PHX_HD float phx_fmaxf(float a, float b) {
    if (b != b)
        return a;
    if (a != a)
        return b;
    return (a > b) ? a : b;
}

PHX_HD float phx_fminf(float a, float b) {
    if (b != b)
        return a;
    if (a != a)
        return b;
    return (a < b) ? a : b;
}

/* OpenCL clamp(x, lo, hi) = fmin(fmax(x, lo), hi). */
PHX_HD float phx_clampf(float x, float lo, float hi) { return phx_fminf(phx_fmaxf(x, lo), hi); }

/* float → u32 with the GPU conversion's saturation: NaN and x ≤ 0 give 0,
 * x ≥ 2^32 gives 0xFFFFFFFF, otherwise truncation toward zero. */
PHX_HD uint32_t phx_f2u(float x) {
    if (!(x > 0.0f))
        return 0u;
    if (x >= 4294967296.0f)
        return 0xFFFFFFFFu;
    return (uint32_t)x;
}

/* 2^e exactly (0 below the smallest denormal, +inf above FLT_MAX). */
PHX_HD float phx_ldexp1(int e) {
    if (e > 127)
        return phx_as_float(0x7f800000u);
    if (e >= -126)
        return phx_as_float((uint32_t)(e + 127) << 23);
    if (e >= -149)
        return phx_as_float(1u << (e + 149));
    return 0.0f;
}

/* Exponent e of frexp(m) for finite m > 0: m = f·2^e with f ∈ [0.5, 1). */
PHX_HD int phx_frexp_exp(float m) {
    uint32_t u = phx_as_uint(m) & 0x7fffffffu;
    if (u < 0x00800000u) {
        u = phx_as_uint(m * 8388608.0f) & 0x7fffffffu; /* ·2^23, exact */
        return (int)((u >> 23) & 0xffu) - 126 - 23;
    }
    return (int)((u >> 23) & 0xffu) - 126;
}

/* ---- sin / cos ---- */

#define PHX_TWO_OVER_PI 0.636619772f
#define PHX_PIO2_1 1.5703125f            /* 0x1.92p+0, 11 bits */
#define PHX_PIO2_2 4.837512969970703e-4f /* 0x1.fb4p-12, 11 bits */
#define PHX_PIO2_3 7.549790126404332e-8f /* 0x1.4442d2p-24 */

PHX_HD float phx_sin_kernel(float r) {
    float r2 = r * r;
    float p = -1.984126984e-4f + r2 * 2.755731922e-6f; /* -1/5040, 1/362880 */
    p = 8.333333333e-3f + r2 * p;                       /* 1/120 */
    p = -1.666666667e-1f + r2 * p;                      /* -1/6 */
    return r + (r * r2) * p;
}

PHX_HD float phx_cos_kernel(float r) {
    float r2 = r * r;
    float p = 2.480158730e-5f + r2 * -2.755731922e-7f; /* 1/40320, -1/3628800 */
    p = -1.388888889e-3f + r2 * p;                      /* -1/720 */
    p = 4.166666667e-2f + r2 * p;                       /* 1/24 */
    return (1.0f - 0.5f * r2) + (r2 * r2) * p;
}

PHX_HD void phx_sincosf(float x, float *s_out, float *c_out) {
    if (x != x || phx_fabsf(x) == phx_as_float(0x7f800000u)) {
        *s_out = phx_as_float(0x7fc00000u);
        *c_out = phx_as_float(0x7fc00000u);
        return;
    }
    float k = phx_floorf(x * PHX_TWO_OVER_PI + 0.5f);
    float r = ((x - k * PHX_PIO2_1) - k * PHX_PIO2_2) - k * PHX_PIO2_3;
    int q = (int)k;
    float s = phx_sin_kernel(r);
    float c = phx_cos_kernel(r);
    switch (q & 3) {
    case 0:
        *s_out = s;
        *c_out = c;
        break;
    case 1:
        *s_out = c;
        *c_out = -s;
        break;
    case 2:
        *s_out = -s;
        *c_out = -c;
        break;
    default:
        *s_out = -c;
        *c_out = s;
        break;
    }
}

PHX_HD float phx_sinf(float x) {
    float s, c;
    phx_sincosf(x, &s, &c);
    return s;
}

PHX_HD float phx_cosf(float x) {
    float s, c;
    phx_sincosf(x, &s, &c);
    return c;
}

/* ---- exp / log / pow ---- */

#define PHX_LOG2E 1.44269504f
#define PHX_LN2_1 0.693115234375f         /* 0x1.62ep-1, 12 bits */
#define PHX_LN2_2 3.194618329871446e-05f  /* 0x1.0bfbe8p-15 */

/* exp(x) = 2^k · p(r), k = nearest integer to x/ln2, r = x − k·ln2 (two-part
 * ln2, each step one fused multiply-add), p the degree-7 Taylor polynomial in
 * Horner form with fused multiply-adds (one rounding per step; an explicit
 * fma is IEEE on both the host and gfx950, so both give the same bits);
 * returns k, p in *p_out. */
PHX_HD int phx_exp_reduce(float x, float *p_out) {
    float k = phx_floorf(__builtin_fmaf(x, PHX_LOG2E, 0.5f));
    float r = __builtin_fmaf(-k, PHX_LN2_2, __builtin_fmaf(-k, PHX_LN2_1, x));
    float p = 1.984126984e-4f;                  /* 1/5040 */
    p = __builtin_fmaf(r, p, 1.388888889e-3f); /* 1/720 */
    p = __builtin_fmaf(r, p, 8.333333333e-3f); /* 1/120 */
    p = __builtin_fmaf(r, p, 4.166666667e-2f); /* 1/24 */
    p = __builtin_fmaf(r, p, 1.666666667e-1f); /* 1/6 */
    p = __builtin_fmaf(r, p, 0.5f);
    p = __builtin_fmaf(r, p, 1.0f);
    p = __builtin_fmaf(r, p, 1.0f);
    *p_out = p;
    return (int)k;
}

PHX_HD float phx_expf(float x) {
    if (x != x)
        return x;
    if (x > 88.72283935f)
        return phx_as_float(0x7f800000u);
    if (x < -103.97208405f)
        return 0.0f;
    float p;
    int ki = phx_exp_reduce(x, &p);
    if (ki > 127)
        return (p * 2.0f) * phx_ldexp1(ki - 1);
    if (ki < -126)
        return (p * phx_ldexp1(ki + 64)) * 5.421010862427522e-20f; /* ·2^-64, one rounding */
    return p * phx_ldexp1(ki);
}

/* a / b correctly rounded from y = RN(1/b) with fused multiply-adds
 * (Markstein 1990): q0 = RN(a·y) lies within 1.5 ulp of a/b; one residual
 * step r = a − b·q (exact in an fma when q is that close and nothing
 * underflows), q1 = RN(q0 + r·y), brings it within one ulp; a second step is
 * then correctly rounded by Markstein's theorem (y within half an ulp of 1/b,
 * q1 faithful).  Valid for normal b and |a| >= 2^-100, where the residuals
 * stay normal; callers handle smaller numerators. */
PHX_HD float phx_div_recip(float a, float b, float y) {
    float q = a * y;
    float r = __builtin_fmaf(-q, b, a);
    q = __builtin_fmaf(r, y, q);
    r = __builtin_fmaf(-q, b, a);
    return __builtin_fmaf(r, y, q);
}

/* phx_expf for -87 <= x <= 88 (2^k a normal float): the same operations
 * without the range branches, so the same bits. */
PHX_HD float phx_expf_normal(float x) {
    float p;
    int ki = phx_exp_reduce(x, &p);
    return p * phx_as_float((uint32_t)(ki + 127) << 23);
}

PHX_HD float phx_logf(float x) {
    if (x != x)
        return x;
    if (x < 0.0f)
        return phx_as_float(0x7fc00000u);
    if (x == 0.0f)
        return phx_as_float(0xff800000u);
    if (x == phx_as_float(0x7f800000u))
        return x;
    uint32_t u = phx_as_uint(x);
    int e = 0;
    if (u < 0x00800000u) {
        u = phx_as_uint(x * 8388608.0f);
        e = -23;
    }
    e += (int)(u >> 23) - 127;
    float m = phx_as_float((u & 0x007fffffu) | 0x3f800000u);
    if (m > 1.414213538f) {
        m = m * 0.5f;
        e += 1;
    }
    float f = m - 1.0f;
    float s = f / (2.0f + f);
    float s2 = s * s;
    float p = 9.090909091e-2f;        /* 1/11 */
    p = 1.111111111e-1f + s2 * p;     /* 1/9 */
    p = 1.428571429e-1f + s2 * p;     /* 1/7 */
    p = 2.0e-1f + s2 * p;             /* 1/5 */
    p = 3.333333333e-1f + s2 * p;     /* 1/3 */
    float two_s = 2.0f * s;
    float logm = two_s + two_s * (s2 * p);
    float fe = (float)e;
    return fe * PHX_LN2_1 + (logm + fe * PHX_LN2_2);
}

PHX_HD float phx_powf(float x, float y) {
    if (y == 0.0f || x == 1.0f)
        return 1.0f;
    if (x != x || y != y)
        return phx_as_float(0x7fc00000u);
    if (x == 0.0f)
        return (y > 0.0f) ? 0.0f : phx_as_float(0x7f800000u);
    if (x < 0.0f)
        return phx_as_float(0x7fc00000u);
    return phx_expf(y * phx_logf(x));
}

/* ---- atan / atan2 / acos ---- */

#define PHX_PI_F 3.14159274f    /* π rounded to float */
#define PHX_PIO2_F 1.57079637f  /* π/2 rounded to float */
#define PHX_PIO6_F 0.523598790f /* π/6 rounded to float */
#define PHX_SQRT3_F 1.73205078f
#define PHX_TAN_PI_12 0.267949194f

PHX_HD float phx_atan_kernel(float t) {
    float t2 = t * t;
    /* Horner for t·(1 − t²/3 + t⁴/5 − t⁶/7 + t⁸/9 − t¹⁰/11 + t¹²/13 − t¹⁴/15) */
    float q = -6.666666667e-2f;     /* -1/15 */
    q = 7.692307692e-2f + t2 * q;   /* 1/13 */
    q = -9.090909091e-2f + t2 * q;  /* -1/11 */
    q = 1.111111111e-1f + t2 * q;   /* 1/9 */
    q = -1.428571429e-1f + t2 * q;  /* -1/7 */
    q = 2.0e-1f + t2 * q;           /* 1/5 */
    q = -3.333333333e-1f + t2 * q;  /* -1/3 */
    return t + (t * t2) * q;
}

PHX_HD float phx_atanf(float x) {
    if (x != x)
        return x;
    float a = phx_fabsf(x);
    int invert = a > 1.0f;
    if (invert)
        a = 1.0f / a;
    float add = 0.0f;
    if (a > PHX_TAN_PI_12) {
        a = (a * PHX_SQRT3_F - 1.0f) / (a + PHX_SQRT3_F);
        add = PHX_PIO6_F;
    }
    float r = add + phx_atan_kernel(a);
    if (invert)
        r = PHX_PIO2_F - r;
    return (x < 0.0f) ? -r : r;
}

PHX_HD float phx_atan2f(float y, float x) {
    if (x != x || y != y)
        return phx_as_float(0x7fc00000u);
    int y_neg = (phx_as_uint(y) >> 31) != 0u;
    int x_neg = (phx_as_uint(x) >> 31) != 0u;
    float ay = phx_fabsf(y), ax = phx_fabsf(x);
    float t;
    if (ay == 0.0f && ax == 0.0f)
        t = 0.0f;
    else if (ay <= ax)
        t = phx_atanf(ay / ax);
    else
        t = PHX_PIO2_F - phx_atanf(ax / ay);
    if (x_neg)
        t = PHX_PI_F - t;
    return y_neg ? -t : t;
}

PHX_HD float phx_acosf(float x) {
    if (x != x || x > 1.0f || x < -1.0f)
        return phx_as_float(0x7fc00000u);
    return 2.0f * phx_atanf(phx_sqrtf((1.0f - x) / (1.0f + x)));
}

#endif /* PHX_MATH_H */
