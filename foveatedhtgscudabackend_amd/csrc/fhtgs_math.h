// Scalar float32 math shared, as source, by the gfx950 kernels (hipcc) and by
// host code (g++). Every routine here uses only IEEE-754 single-precision
// operations that both compilers implement with a single correct rounding:
// + - * /, sqrtf, fmaf, floorf/ceilf/rintf, comparisons and bit casts. Both
// sides are built with -ffp-contract=off, so no multiply-add is fused unless
// it is written as fmaf() here or at the call site. The same input therefore
// gives the same bits on the CPU and on the GPU.
//
// Transcendentals are our own (no vendor libm), from textbook methods:
//   fh_expf : Cody-Waite reduction x = k ln2 + r with a two-part ln2, then the
//             degree-7 Taylor polynomial of e^r in Horner form, then exact 2^k
//             scaling (W. J. Cody, W. Waite, "Software Manual for the
//             Elementary Functions", 1980, ch. 6).
//   fh_logf : x = 2^e m with m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s),
//             s = (m-1)/(m+1), odd series of atanh to s^9 (same reference, ch. 5).
//   fh_atanf/fh_atan2f, fh_asinf, fh_tanf, fh_tanhf: see each function.
// Accuracy of each is checked against libm in the oracle's self tests.
#pragma once

#include <stdint.h>
#include <math.h>

#if defined(__HIPCC__) || defined(__HIP__)
#define FH_HD __host__ __device__ __forceinline__
#else
#define FH_HD static inline
#endif

FH_HD uint32_t fh_as_u32(float x) {
    union { float f; uint32_t u; } c;
    c.f = x;
    return c.u;
}

FH_HD float fh_as_f32(uint32_t x) {
    union { float f; uint32_t u; } c;
    c.u = x;
    return c.f;
}

FH_HD float fh_fma(float a, float b, float c) { return fmaf(a, b, c); }

// IEEE minNum/maxNum: a NaN operand yields the other operand.
FH_HD float fh_fmin(float a, float b) { return fminf(a, b); }
FH_HD float fh_fmax(float a, float b) { return fmaxf(a, b); }

// clamp(f, a, b) = max(a, min(f, b)) (the reference's vector-math clamp order)
FH_HD float fh_clamp(float f, float a, float b) { return fh_fmax(a, fh_fmin(f, b)); }

// Clamp to [0, 1]; NaN maps to 0.
FH_HD float fh_saturate(float x) { return x > 0.0f ? (x < 1.0f ? x : 1.0f) : 0.0f; }

// float -> int32, rounding toward -inf / +inf, saturating, NaN -> 0.
FH_HD int32_t fh_f2i_rd(float x) {
    const float f = floorf(x);
    if (!(f == f)) return 0;
    if (f >= 2147483648.0f) return 2147483647;
    if (f < -2147483648.0f) return (-2147483647 - 1);
    return (int32_t)f;
}

FH_HD int32_t fh_f2i_ru(float x) {
    const float f = ceilf(x);
    if (!(f == f)) return 0;
    if (f >= 2147483648.0f) return 2147483647;
    if (f < -2147483648.0f) return (-2147483647 - 1);
    return (int32_t)f;
}

// 2^k for k in [-126, 127] as a normal float.
FH_HD float fh_pow2i(int32_t k) { return fh_as_f32((uint32_t)(k + 127) << 23); }

// Exact-then-once-rounded scaling p * 2^k for k in [-150, 128], p in [0.5, 2].
FH_HD float fh_ldexp_small(float p, int32_t k) {
    const int32_t a = k / 2;
    const int32_t b = k - a;
    return (p * fh_pow2i(a)) * fh_pow2i(b);
}

FH_HD float fh_expf(float x) {
    if (!(x == x)) return x;
    if (x > 0x1.62e430p+6f) return fh_as_f32(0x7f800000u);   // > ln(FLT_MAX)
    if (x < -0x1.9fe368p+6f) return 0.0f;                     // < ln(2^-150)
    const float k = rintf(x * 0x1.715476p+0f);                // round(x / ln2)
    float r = fh_fma(-k, 0x1.62e400p-1f, x);                  // ln2 high part (exact k*hi)
    r = fh_fma(-k, 0x1.7f7d1cp-20f, r);                       // ln2 low part
    float p = 0x1.a01a02p-13f;                                // 1/7!
    p = fh_fma(p, r, 0x1.6c16c2p-10f);                        // 1/6!
    p = fh_fma(p, r, 0x1.111112p-7f);                         // 1/5!
    p = fh_fma(p, r, 0x1.555556p-5f);                         // 1/4!
    p = fh_fma(p, r, 0x1.555556p-3f);                         // 1/3!
    p = fh_fma(p, r, 0.5f);
    p = fh_fma(p, r, 1.0f);
    p = fh_fma(p, r, 1.0f);
    return fh_ldexp_small(p, (int32_t)k);
}

// The reduction and polynomial of fh_expf: e^x = p * 2^k, k = rint(x / ln2)
// (the same operations as fh_expf, without its range checks).
FH_HD float fh_expf_core(float x, float* k_out) {
    const float k = rintf(x * 0x1.715476p+0f);
    float r = fh_fma(-k, 0x1.62e400p-1f, x);
    r = fh_fma(-k, 0x1.7f7d1cp-20f, r);
    float p = 0x1.a01a02p-13f;
    p = fh_fma(p, r, 0x1.6c16c2p-10f);
    p = fh_fma(p, r, 0x1.111112p-7f);
    p = fh_fma(p, r, 0x1.555556p-5f);
    p = fh_fma(p, r, 0x1.555556p-3f);
    p = fh_fma(p, r, 0.5f);
    p = fh_fma(p, r, 1.0f);
    p = fh_fma(p, r, 1.0f);
    *k_out = k;
    return p;
}

// fh_expf for x in [-64, 0] or NaN, without the range checks: the same
// operations, and the final 2^k scaling (exact in both forms: p is within
// [0.7, 1.42] and k >= -93, so p * 2^k is a normal float) done in one
// multiply, so the result is bit-identical to fh_expf there. The blend's
// exponent -rho^2/2 is >= -5.55 (rho^2 <= 11.0825) or NaN.
FH_HD float fh_expf_m64(float x) {
    float k;
    const float p = fh_expf_core(x, &k);
    const int32_t ki = (k == k) ? (int32_t)k : 0;
    return p * fh_pow2i(ki);
}

FH_HD float fh_expf_nonpos(float x) {
    if (x < -64.0f) return fh_expf(x);
    return fh_expf_m64(x);
}

FH_HD float fh_logf(float x) {
    if (!(x == x)) return x;
    if (x < 0.0f) return fh_as_f32(0x7fc00000u);
    if (x == 0.0f) return fh_as_f32(0xff800000u);
    if (x == fh_as_f32(0x7f800000u)) return x;
    uint32_t bits = fh_as_u32(x);
    int32_t e_adj = 0;
    if (bits < 0x00800000u) {                                 // subnormal: scale by 2^23
        bits = fh_as_u32(x * 8388608.0f);
        e_adj = -23;
    }
    int32_t e = (int32_t)((bits >> 23) & 0xffu) - 127 + e_adj;
    float m = fh_as_f32((bits & 0x007fffffu) | 0x3f800000u);  // [1, 2)
    if (m > 0x1.6a09e6p+0f) {                                 // > sqrt(2)
        m = m * 0.5f;
        e += 1;
    }
    const float f = m - 1.0f;                                  // exact
    const float s = f / (2.0f + f);
    const float z = s * s;
    float q = fh_fma(z, 0x1.c71c72p-3f, 0x1.24924ap-2f);       // 2/9, 2/7
    q = fh_fma(z, q, 0x1.99999ap-2f);                          // 2/5
    q = fh_fma(z, q, 0x1.555556p-1f);                          // 2/3
    q = q * z;
    // log m = 2s + s*q, written as f - s*(f - q) to keep the leading term exact-ish
    const float log_m = fh_fma(-s, f - q, f);
    const float ef = (float)e;
    return fh_fma(ef, 0x1.62e400p-1f, fh_fma(ef, 0x1.7f7d1cp-20f, log_m));
}

// atan on [0, inf): reduce to |t| <= tan(pi/12) with the identities
// atan(x) = pi/2 - atan(1/x) and atan(x) = pi/6 + atan((x*sqrt3 - 1)/(x + sqrt3)),
// then the odd Taylor series of atan to t^13 (Cody & Waite, ch. 11).
FH_HD float fh_atan_pos(float x) {
    bool inv = false;
    if (x > 1.0f) { x = 1.0f / x; inv = true; }
    bool shift = false;
    if (x > 0x1.126146p-2f) {                                 // tan(pi/12) = 2 - sqrt(3)
        x = (fh_fma(x, 0x1.bb67aep+0f, -1.0f)) / (x + 0x1.bb67aep+0f);
        shift = true;
    }
    const float z = x * x;
    float p = -0x1.3b13b2p-4f;                                // -1/13
    p = fh_fma(p, z, 0x1.745d18p-4f);                         //  1/11
    p = fh_fma(p, z, -0x1.c71c72p-4f);                        // -1/9
    p = fh_fma(p, z, 0x1.24924ap-3f);                         //  1/7
    p = fh_fma(p, z, -0x1.99999ap-3f);                        // -1/5
    p = fh_fma(p, z, 0x1.555556p-2f);                         //  1/3
    float r = fh_fma(-(p * z), x, x);                         // x - x^3/3 + ...
    if (shift) r = r + 0x1.0c1524p-1f;                        // pi/6
    if (inv) r = 0x1.921fb6p+0f - r;                          // pi/2
    return r;
}

FH_HD float fh_atanf(float x) {
    if (!(x == x)) return x;
    return x < 0.0f ? -fh_atan_pos(-x) : fh_atan_pos(x);
}

FH_HD float fh_atan2f(float y, float x) {
    const float pi = 0x1.921fb6p+1f;
    if (!(x == x) || !(y == y)) return x + y;
    if (x == 0.0f && y == 0.0f) {
        const bool xneg = (fh_as_u32(x) >> 31) != 0u;
        const float r = xneg ? pi : 0.0f;
        return (fh_as_u32(y) >> 31) ? -r : r;
    }
    const float ax = fabsf(x), ay = fabsf(y);
    float r;
    if (ax >= ay) r = fh_atan_pos(ay / ax);
    else r = 0x1.921fb6p+0f - fh_atan_pos(ax / ay);
    if (x < 0.0f) r = pi - r;
    return y < 0.0f ? -r : r;
}

// asin(x) = atan2(x, sqrt((1-x)(1+x)))
FH_HD float fh_asinf(float x) {
    if (!(x == x) || x > 1.0f || x < -1.0f) return fh_as_f32(0x7fc00000u);
    return fh_atan2f(x, sqrtf((1.0f - x) * (1.0f + x)));
}

// tan(x): x = k pi/2 + r (three-part pi/2), sin and cos of r by Taylor series,
// tan = sin/cos, or -cos/sin for odd k. For |x| up to a few pi (used on angles).
FH_HD float fh_tanf(float x) {
    if (!(x == x) || fabsf(x) == fh_as_f32(0x7f800000u)) return fh_as_f32(0x7fc00000u);
    const float k = rintf(x * 0x1.45f306p-1f);                // x * 2/pi
    float r = fh_fma(-k, 0x1.921f00p+0f, x);
    r = fh_fma(-k, 0x1.6a8880p-17f, r);
    r = fh_fma(-k, 0x1.68c000p-39f, r);
    const float z = r * r;
    float sp = 0x1.71de3ap-19f;                               //  1/9!
    sp = fh_fma(sp, z, -0x1.a01a02p-13f);                     // -1/7!
    sp = fh_fma(sp, z, 0x1.111112p-7f);                       //  1/5!
    sp = fh_fma(sp, z, -0x1.555556p-3f);                      // -1/3!
    const float s = fh_fma(sp * z, r, r);
    float cp = -0x1.27e4fcp-22f;                              // -1/10!
    cp = fh_fma(cp, z, 0x1.a01a02p-16f);                      //  1/8!
    cp = fh_fma(cp, z, -0x1.6c16c2p-10f);                     // -1/6!
    cp = fh_fma(cp, z, 0x1.555556p-5f);                       //  1/4!
    cp = fh_fma(cp, z, -0.5f);
    const float c = fh_fma(cp, z, 1.0f);
    const int32_t ki = (int32_t)k;
    return (ki & 1) ? -(c / s) : (s / c);
}

// tanh(x) = sign(x) (1 - 2/(e^{2|x|} + 1)); odd Taylor series for |x| < 2^-4.
FH_HD float fh_tanhf(float x) {
    if (!(x == x)) return x;
    const float ax = fabsf(x);
    float r;
    if (ax > 9.0f) {
        r = 1.0f;
    } else if (ax < 0.0625f) {
        const float z = x * x;
        float p = 0x1.ba1ba2p-5f;                             //  17/315 (x^7 term)
        p = fh_fma(p, z, -0x1.111112p-3f);                    // -2/15
        p = fh_fma(p, z, 0x1.555556p-2f);                     //  1/3
        return fh_fma(-(p * z), x, x);
    } else {
        r = 1.0f - 2.0f / (fh_expf(2.0f * ax) + 1.0f);
    }
    return x < 0.0f ? -r : r;
}
