/*
 * athx_detmath.h -- the arithmetic contract of the splat projection and the
 * blend's per-pair decisions.
 *
 * The rasterizer's sorted keys, tile rectangles and tile counts must come out
 * bit-identical on the device and in a host restatement. They depend on FP32
 * exp (log-scales, the opacity sigmoid), log (the cut radius), sqrt and
 * division. IEEE sqrt and division are correctly rounded on both sides (hipcc
 * keeps fp32 divide/sqrt correctly rounded by default; glibc's are too); exp
 * and log are not shared by any two libraries, so they are defined here, once,
 * from published methods, using only operations that are correctly rounded
 * everywhere (fmaf, +, *, rint and exact power-of-two scaling):
 *
 *   exp: Cody-Waite two-constant reduction x = k ln2 + r, |r| <= ln2/2, and the
 *        degree-6 minimax polynomial of the Cephes library's expf
 *        (S. L. Moshier, Cephes Math Library, 1984-1992), evaluated by Horner
 *        with fused multiply-adds; the result is scaled by 2^k built from bits.
 *   log: frexp-style reduction to m in [sqrt(1/2), sqrt(2)), and the degree-9
 *        minimax polynomial of Cephes' logf, with the same two-constant ln2.
 *
 * Both stay within 2 ulp of the true value over the ranges the rasterizer
 * uses (checked against double-precision libm by the test suite).
 *
 * Every translation unit that includes this header and wants bit-identical
 * results must be compiled with floating-point contraction off
 * (-ffp-contract=off): a*b+c written in the source is then a multiply and an
 * add on both sides, and only the explicit fmaf() calls below fuse.
 *
 * The header is plain C99 so the host restatement (a C file) and the HIP
 * kernels include the same text.
 */
#ifndef ATHX_DETMATH_H
#define ATHX_DETMATH_H

#include <stdint.h>

#if defined(__HIPCC__)
#define ATHX_DM_FN static inline __host__ __device__
#else
#include <math.h>
#include <string.h>
#define ATHX_DM_FN static inline
#endif

ATHX_DM_FN uint32_t athx_dm_bits(float f) {
#if defined(__HIPCC__)
    return __builtin_bit_cast(uint32_t, f);
#else
    uint32_t u;
    memcpy(&u, &f, sizeof u);
    return u;
#endif
}

ATHX_DM_FN float athx_dm_float(uint32_t u) {
#if defined(__HIPCC__)
    return __builtin_bit_cast(float, u);
#else
    float f;
    memcpy(&f, &u, sizeof f);
    return f;
#endif
}

/* e^r for x = k ln2 + r: the reduction and the polynomial. */
ATHX_DM_FN float athx_dm_exp_reduced(float x, float k) {
    /* ln2 = C1 + C2; C1 has 9 significant bits so k*C1 is exact for |k| <= 128. */
    float r = fmaf(k, -0.693359375f, x);
    r = fmaf(k, 2.12194440e-4f, r);
    float p = 1.9875691500e-4f;
    p = fmaf(p, r, 1.3981999507e-3f);
    p = fmaf(p, r, 8.3334519073e-3f);
    p = fmaf(p, r, 4.1665795894e-2f);
    p = fmaf(p, r, 1.6666665459e-1f);
    p = fmaf(p, r, 5.0000001201e-1f);
    const float r2 = r * r;
    p = fmaf(p, r2, r);
    return p + 1.0f;
}

/* e^x for x in [-87, 88]; 0 below, +inf above. */
ATHX_DM_FN float athx_expf(float x) {
    if (!(x > -87.0f)) {
        return x == x ? 0.0f : x; /* NaN passes through */
    }
    if (x > 88.0f) {
        return athx_dm_float(0x7F800000u);
    }
    const float k = rintf(x * 1.44269504088896341f);
    const float p = athx_dm_exp_reduced(x, k);
    /* 2^k, k in [-126, 127]: exact, and the product is exact unless it is subnormal. */
    const int ki = (int)k;
    if (ki < -125) {
        /* Two exact halvings keep the scale a normal number. */
        return (p * athx_dm_float((uint32_t)(ki + 127 + 2) << 23)) * 0.25f;
    }
    return p * athx_dm_float((uint32_t)(ki + 127) << 23);
}

/* athx_expf without its range checks, for callers whose x is in [-80, 80]:
 * there 2^k is a normal number and the result is athx_expf's. */
ATHX_DM_FN float athx_expf_core(float x) {
    const float k = rintf(x * 1.44269504088896341f);
    return athx_dm_exp_reduced(x, k) * athx_dm_float((uint32_t)((int)k + 127) << 23);
}

/* ln x for positive normal x; -inf at 0, NaN below 0, x at +inf. */
ATHX_DM_FN float athx_logf(float x) {
    if (!(x > 0.0f)) {
        return x == 0.0f ? athx_dm_float(0xFF800000u) : athx_dm_float(0x7FC00000u);
    }
    uint32_t u = athx_dm_bits(x);
    if (u >= 0x7F800000u) {
        return x;
    }
    int e;
    if (u < 0x00800000u) {
        /* Subnormal: scale into the normal range first. */
        u = athx_dm_bits(x * 16777216.0f);
        e = (int)(u >> 23) - 126 - 24;
    } else {
        e = (int)(u >> 23) - 126;
    }
    float m = athx_dm_float((u & 0x007FFFFFu) | 0x3F000000u); /* [0.5, 1) */
    if (m < 0.707106781186547524f) {
        e -= 1;
        m = (m + m) - 1.0f;
    } else {
        m = m - 1.0f;
    }
    const float z = m * m;
    float y = 7.0376836292e-2f;
    y = fmaf(y, m, -1.1514610310e-1f);
    y = fmaf(y, m, 1.1676998740e-1f);
    y = fmaf(y, m, -1.2420140846e-1f);
    y = fmaf(y, m, 1.4249322787e-1f);
    y = fmaf(y, m, -1.6668057665e-1f);
    y = fmaf(y, m, 2.0000714765e-1f);
    y = fmaf(y, m, -2.4999993993e-1f);
    y = fmaf(y, m, 3.3333331174e-1f);
    y = (y * m) * z;
    const float fe = (float)e;
    y = fmaf(fe, -2.12194440e-4f, y);
    y = fmaf(z, -0.5f, y);
    float r = m + y;
    r = fmaf(fe, 0.693359375f, r);
    return r;
}

/* 1 / (1 + e^-x): the opacity a logit stands for. */
ATHX_DM_FN float athx_sigmoidf(float x) {
    return 1.0f / (1.0f + athx_expf(-x));
}

/* x^y for x > 0 (the sRGB curve's exponent): e^(y ln x). */
ATHX_DM_FN float athx_powf_pos(float x, float y) {
    return athx_expf(y * athx_logf(x));
}

#endif /* ATHX_DETMATH_H */
