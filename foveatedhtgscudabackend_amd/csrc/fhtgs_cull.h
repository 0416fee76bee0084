// Per-tile culling of Gaussians in the blend, shared by the gfx950 kernel
// and host-side audits (plain C++ with -ffp-contract=off sees the same
// float32 operations).
//
// A Gaussian's pixel alpha is opacity * exp(-rho^2 / 2) with rho^2 the
// squared distance of the pixel ray from the Gaussian centre in its unit
// frame (inference.cuh:202-222), so alpha >= 1/255 only inside the screen
// conic rho^2 <= c, c = 2 ln(255 opacity). With the rows a1, a2, a4 of VPM*T
// (VPMT1, VPMT2, VPMT4) and D = diag(c, c, c, -1), the dual of that conic is
//     C* = sum_k f_k a_i,k a_j,k,  f = D / s,  s = sum_k D_k a4_k^2
// (quadric silhouette bounds, as in preprocess.cuh's bounding box). When
// s < 0 and a4.w > 0 the conic is an ellipse with centre
// (cx, cy) = (C*_14, C*_24) and shape matrix
//     M = [[cx^2 - C*_11, cx cy - C*_12], [cx cy - C*_12, cy^2 - C*_22]],
// i.e. the points p with (p - c)^T M^-1 (p - c) <= 1; sqrt(M11), sqrt(M22)
// are its half extents. The test below keeps a Gaussian iff that ellipse,
// grown by kCullGrowPx (M + g^2 I), meets the sample rectangle grown by
// kCullMarginPx: first the box test, then the exact minimum of the convex
// quadratic over the rectangle (centre inside, else the clamped minimum on
// each edge). Anything that is not a proper ellipse is kept. Dropping a
// Gaussian that reaches no sample of the tile leaves the image unchanged.
#pragma once
#include "fhtgs_math.h"

constexpr float kCullMarginPx = 0.0625f;
constexpr float kCullGrowPx = 0.0625f;

FH_HD bool fh_tile_may_touch(const float a1[4], const float a2[4], const float a4[4], float opacity,
                             float x0, float x1, float y0, float y1) {
    const float c = 2.0f * fh_logf(opacity * 255.0f);
    const float s = fmaf(c, fmaf(a4[2], a4[2], fmaf(a4[1], a4[1], a4[0] * a4[0])), -(a4[3] * a4[3]));
    if (!(s < 0.0f) || !(a4[3] > 0.0f)) return true;
    const float rs = 1.0f / s;
    const float fc = rs * c, fw = -rs;
    // Screen coordinates relative to the rectangle centre (ox, oy): the rows
    // become a1 - ox a4 and a2 - oy a4. Near the tile the centre offset and
    // the shape matrix are then small numbers instead of differences of
    // squares of pixel coordinates in the thousands.
    const float ox = 0.5f * (x0 + x1), oy = 0.5f * (y0 + y1);
    const float hx = 0.5f * (x1 - x0) + kCullMarginPx, hy = 0.5f * (y1 - y0) + kCullMarginPx;
    float b1[4], b2[4];
    for (int k = 0; k < 4; ++k) {
        b1[k] = fmaf(-a4[k], ox, a1[k]);
        b2[k] = fmaf(-a4[k], oy, a2[k]);
    }
    // sum_k f_k u_k v_k with f = (fc, fc, fc, fw)
    auto dual = [&](const float* u, const float* v) {
        return fmaf(fw, u[3] * v[3], fc * fmaf(u[2], v[2], fmaf(u[1], v[1], u[0] * v[0])));
    };
    const float cx = dual(b1, a4), cy = dual(b2, a4);
    const float g2 = kCullGrowPx * kCullGrowPx;
    const float m11 = fmaf(cx, cx, -dual(b1, b1)) + g2;
    const float m22 = fmaf(cy, cy, -dual(b2, b2)) + g2;
    const float m12 = fmaf(cx, cy, -dual(b1, b2));
    const float ex = sqrtf(fh_fmax(m11, 0.0f)), ey = sqrtf(fh_fmax(m22, 0.0f));
    if ((cx + ex < -hx) || (cx - ex > hx) || (cy + ey < -hy) || (cy - ey > hy)) return false;
    if (cx >= -hx && cx <= hx && cy >= -hy && cy <= hy) return true;
    const float det = fmaf(m11, m22, -(m12 * m12));
    if (!(det > 0.0f) || !(m11 > 0.0f) || !(m22 > 0.0f)) return true;
    // q(u, v) = m22 u^2 - 2 m12 u v + m11 v^2 <= det  <=>  inside the ellipse
    auto q = [&](float u, float v) { return fmaf(m22, u * u, fmaf(m11, v * v, -2.0f * (m12 * (u * v)))); };
    const float u0 = -hx - cx, u1 = hx - cx, v0 = -hy - cy, v1 = hy - cy;
    const float ru = m12 / m11, rv = m12 / m22;
    // edges x = const: the minimum over v is at v = (m12 / m11) u
    float best = q(u0, fh_clamp(ru * u0, v0, v1));
    best = fh_fmin(best, q(u1, fh_clamp(ru * u1, v0, v1)));
    // edges y = const: the minimum over u is at u = (m12 / m22) v
    best = fh_fmin(best, q(fh_clamp(rv * v0, u0, u1), v0));
    best = fh_fmin(best, q(fh_clamp(rv * v1, u0, u1), v1));
    return best <= det;
}

// Pixel-level pre-test of the blend's first pass, a superset of the pixels a
// Gaussian can reach: the support ellipse of the same dual conic at the
// level c' = c (1 + kPixelLevelRel) + kPixelLevelAbs, with c the smaller of
// 2 ln(255 opacity) and the reference's cutoff 11.0825 (every pixel the
// reference keeps has rho^2 <= c up to rounding), its shape matrix grown by
// kPixelGrowPx^2 I, in coordinates relative to the tile centre (ox, oy).
// k[0..5] = ellipse centre (cx, cy), the quadratic form q(du, dv) =
// k2 du^2 + k3 du dv + k4 dv^2 (k2 = M22, k3 = -2 M12, k4 = M11) and its bound
// k5 = det M; a pixel at offset (u, v) passes when q(u - cx, v - cy) <= k5.
// k5 = -1 marks a Gaussian without a proper, finite, not too eccentric
// ellipse: the first pass then uses the exact test. Conservativeness is
// audited per pixel (tests/cull_audit.cpp), as for the tile cull.
constexpr float kPixelLevelRel = 0.02f;
constexpr float kPixelLevelAbs = 0.02f;
constexpr float kPixelGrowPx = 0.25f;

FH_HD void fh_pixel_conic(const float a1[4], const float a2[4], const float a4[4], float opacity, float ox,
                          float oy, float k[6]) {
    k[0] = k[1] = k[2] = k[3] = k[4] = 0.0f;
    k[5] = -1.0f;
    float c = 2.0f * fh_logf(opacity * 255.0f);
    c = c < 11.0825270903f ? c : 11.0825270903f;  // NaN: the cutoff
    c = fmaf(c, 1.0f + kPixelLevelRel, kPixelLevelAbs);
    if (!(c > 0.0f)) return;
    const float s = fmaf(c, fmaf(a4[2], a4[2], fmaf(a4[1], a4[1], a4[0] * a4[0])), -(a4[3] * a4[3]));
    if (!(s < 0.0f) || !(a4[3] > 0.0f)) return;
    const float rs = 1.0f / s;
    const float fc = rs * c, fw = -rs;
    float b1[4], b2[4];
    for (int j = 0; j < 4; ++j) {
        b1[j] = fmaf(-a4[j], ox, a1[j]);
        b2[j] = fmaf(-a4[j], oy, a2[j]);
    }
    auto dual = [&](const float* u, const float* v) {
        return fmaf(fw, u[3] * v[3], fc * fmaf(u[2], v[2], fmaf(u[1], v[1], u[0] * v[0])));
    };
    const float cx = dual(b1, a4), cy = dual(b2, a4);
    const float g2 = kPixelGrowPx * kPixelGrowPx;
    const float m11 = fmaf(cx, cx, -dual(b1, b1)) + g2;
    const float m22 = fmaf(cy, cy, -dual(b2, b2)) + g2;
    const float m12 = fmaf(cx, cy, -dual(b1, b2));
    const float det = fmaf(m11, m22, -(m12 * m12));
    // finite and small enough that q cannot overflow for |u|, |v| <= 16
    const bool sane = fabsf(cx) < 1e9f && fabsf(cy) < 1e9f && m11 < 1e17f && m22 < 1e17f &&
                      fabsf(m12) < 1e17f;
    if (!sane || !(m11 > 0.0f) || !(m22 > 0.0f) || !(det > 1e-4f * (m11 * m22))) return;
    k[0] = cx;
    k[1] = cy;
    k[2] = m22;
    k[3] = -2.0f * m12;
    k[4] = m11;
    k[5] = det;
}

FH_HD bool fh_pixel_in_conic(const float k[6], float u, float v) {
    const float du = u - k[0], dv = v - k[1];
    return fmaf(k[2], du * du, fmaf(k[4], dv * dv, k[3] * (du * dv))) <= k[5];
}

// Bound on rho^2 = num / den beyond which a pixel cannot pass the
// reference's alpha test (inference.cuh:216-224) for this opacity:
// alpha = min(o exp(-rho^2 / 2), 1) >= 1/255 needs rho^2 <= 2 ln(255 o). The
// margin of 0.01 covers the rounding of the logarithm, the exponential, the
// reciprocal and the products (a few ulp each, below 1e-5 in rho^2), so every
// pair the reference keeps passes `!(num > bound * den)`. From an opacity of
// about 0.995 the bound is the reference's own cutoff constant, and that
// test is then the reference's cutoff test itself. NaN maps to the cutoff
// constant and -inf to 0 (the selects, unlike min/max, are explicit about
// NaN), which keeps every pair whose num or den is NaN.
FH_HD float fh_pixel_cutoff(float opacity) {
    float c = 2.0f * fh_logf(255.0f * opacity) + 0.01f;
    c = c < 11.0825270903f ? c : 11.0825270903f;
    return c > 0.0f ? c : 0.0f;
}
