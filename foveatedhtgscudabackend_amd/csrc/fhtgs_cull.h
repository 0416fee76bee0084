// Per-tile and per-pixel culling of Gaussians in the blend, shared by the
// gfx950 kernels and host-side audits (plain C++ with -ffp-contract=off sees
// the same float32 operations).
//
// A Gaussian's pixel alpha is opacity * exp(-rho^2 / 2) with rho^2 the
// squared distance of the pixel ray from the Gaussian centre in its unit
// frame (inference.cuh:202-222), so alpha >= 1/255 only inside the screen
// conic rho^2 <= c, c = 2 ln(255 opacity); the reference also drops every
// pixel with rho^2 > 11.0825 (its cutoff constant). With the rows a1, a2, a4
// of VPM*T (VPMT1, VPMT2, VPMT4) and D = diag(c, c, c, -1), the dual of that
// conic is
//     C* = sum_k f_k a_i,k a_j,k,  f = D / s,  s = sum_k D_k a4_k^2
// (quadric silhouette bounds, as in preprocess.cuh's bounding box). When
// s < 0 and a4.w > 0 the conic is an ellipse with centre
// (cx, cy) = (C*_14, C*_24) and shape matrix
//     M = [[cx^2 - C*_11, cx cy - C*_12], [cx cy - C*_12, cy^2 - C*_22]],
// i.e. the points p with (p - c)^T M^-1 (p - c) <= 1; sqrt(M11), sqrt(M22)
// are its half extents.
//
// The preprocess stores one such ellipse per visible Gaussian and view (the
// ellipse record, fh_ellipse_record): the support ellipse at the level
// c' = min(c, 11.0825) (1 + kPixelLevelRel) + kPixelLevelAbs, its shape
// matrix grown by kPixelGrowPx^2 I -- a superset of every pixel the
// reference's cutoff and alpha tests keep, up to rounding. The blend then
// only translates it per tile:
//   * fh_mask_prep / fh_mask_row_w2 / fh_mask_row_bits: the listed
//     Gaussian's sample mask, which samples of the tile its ellipse can
//     contain (below). The blend drops a Gaussian whose mask is empty (it
//     reaches no sample of the tile, so the image is unchanged) and its
//     pixels' first pass reads one bit per staged Gaussian.
// A record whose ellipse is not proper and finite is flagged: its mask, as
// for a too eccentric ellipse, holds every sample. Conservativeness is
// audited per (tile, Gaussian) pair and per sample with these functions
// (tests/cull_audit.cpp).
#pragma once
#include "fhtgs_math.h"

constexpr float kPixelLevelRel = 0.02f;
constexpr float kPixelLevelAbs = 0.02f;
constexpr float kPixelGrowPx = 0.25f;
// centres beyond this many pixels from the origin are flagged (their
// float32 position would lose more than 1/256 px)
constexpr float kEllipseMaxCoord = 65536.0f;

// The ellipse record, 5 floats, stored as a float4 and a float:
//   r[0], r[1]  centre (cx, cy) in pixel coordinates
//   r[2], r[3]  m11, m22 of the grown shape matrix M, or m11 = -1 when the
//               ellipse is unusable
//   r[4]        m12
// `c2` is 2 ln(255 opacity) as fh_logf computes it (the preprocess has it
// already for the tile bounds). The shape is computed about an anchor, the
// first estimate of the centre rounded to whole pixels, where the rows
// a1 - gx a4, a2 - gy a4 and the centre offset are small numbers; only the
// centre is then moved back to pixel coordinates (one rounding).
FH_HD void fh_ellipse_record(const float a1[4], const float a2[4], const float a4[4], float c2, float r[5]) {
    r[0] = r[1] = r[3] = r[4] = 0.0f;
    r[2] = -1.0f;
    float c = c2 < 11.0825270903f ? c2 : 11.0825270903f;  // NaN: the cutoff
    c = fmaf(c, 1.0f + kPixelLevelRel, kPixelLevelAbs);
    if (!(c > 0.0f)) return;
    const float s = fmaf(c, fmaf(a4[2], a4[2], fmaf(a4[1], a4[1], a4[0] * a4[0])), -(a4[3] * a4[3]));
    if (!(s < 0.0f) || !(a4[3] > 0.0f)) return;
    const float rs = 1.0f / s;
    const float fc = rs * c, fw = -rs;
    // sum_k f_k u_k v_k with f = (fc, fc, fc, fw)
    auto dual = [&](const float* u, const float* v) {
        return fmaf(fw, u[3] * v[3], fc * fmaf(u[2], v[2], fmaf(u[1], v[1], u[0] * v[0])));
    };
    const float gx0 = dual(a1, a4), gy0 = dual(a2, a4);
    if (!(fabsf(gx0) < kEllipseMaxCoord) || !(fabsf(gy0) < kEllipseMaxCoord)) return;
    const float gx = rintf(gx0), gy = rintf(gy0);
    float b1[4], b2[4];
    for (int j = 0; j < 4; ++j) {
        b1[j] = fmaf(-a4[j], gx, a1[j]);
        b2[j] = fmaf(-a4[j], gy, a2[j]);
    }
    const float dx = dual(b1, a4), dy = dual(b2, a4);
    const float g2 = kPixelGrowPx * kPixelGrowPx;
    const float m11 = fmaf(dx, dx, -dual(b1, b1)) + g2;
    const float m22 = fmaf(dy, dy, -dual(b2, b2)) + g2;
    const float m12 = fmaf(dx, dy, -dual(b1, b2));
    const float det = fmaf(m11, m22, -(m12 * m12));
    const float cx = gx + dx, cy = gy + dy;
    // finite, and small enough that the quadratic forms cannot overflow
    // near the ellipse
    const bool sane = fabsf(cx) < kEllipseMaxCoord && fabsf(cy) < kEllipseMaxCoord && m11 < 1e17f &&
                      m22 < 1e17f && fabsf(m12) < 1e17f;
    if (!sane || !(m11 > 0.0f) || !(m22 > 0.0f) || !(det > 0.0f)) return;
    r[0] = cx;
    r[1] = cy;
    r[2] = m11;
    r[3] = m22;
    r[4] = m12;
}

// ---- sample masks: which samples of a tile an ellipse record can reach
//
// The samples of a tile are (bx + c step, by + r step), c < cols, r < rows
// (8 x 8 at stride 1 or 2, 4 x 4 at stride 4); bit r * cols + c of the mask
// stands for sample (c, r). In row r the record's ellipse spans
//     x = x_c(y) +- w(y),  x_c(y) = cx + (m12 / m22)(y - cy),
//     w(y)^2 = det / m22^2 * (m22 - (y - cy)^2),
// so the row's bits are the columns of [x_c - w, x_c + w] widened by
// kMaskMarginPx on each side. The blend's staging lane computes the mask of
// each listed Gaussian (the square root with the hardware's v_sqrt_f32) and
// stages it only if the mask holds a sample -- that is the blend's tile cull
// -- and the pixels' first pass then reads one bit per staged Gaussian.
// Records that are flagged, or too eccentric for det to be accurate
// (det <= 1e-4 m11 m22), get every sample. tests/cull_audit.cpp checks, with
// the square root lowered by a relative 2^-20, that the masks keep every
// sample the reference's cutoff and alpha tests keep.
//
// For a usable record every quantity below is finite: |cx|, |cy| < 65536,
// m11, m22 >= 1/16 (the growth) and < 1e17, so s < 2e18 and a <= m11 / m22 <
// 2e18; a * (m22 - dy^2) may round to -inf or +inf, which the clamps below
// turn into an empty or a full row.
constexpr float kMaskMarginPx = 0.015625f;

struct FhMaskPrep {
    float u0;     // (cx - bx) / step: the centre in sample columns
    float dy0;    // by - cy: the first row's offset from the centre
    float s;      // m12 / m22 / step: column shift per pixel of dy
    float a;      // det / m22^2 / step^2: w(y)^2 in columns^2 is a (m22 - dy^2)
    float m22;
    bool all;     // no usable ellipse: every sample
};

// inv22 and inv_step: 1 / m22 and 1 / step, correctly rounded (the kernels
// use a reciprocal that equals IEEE division on every float).
FH_HD FhMaskPrep fh_mask_prep(const float r[5], float inv22, float bx, float by, float inv_step) {
    FhMaskPrep p;
    const float m11 = r[2], m22 = r[3], m12 = r[4];
    const float det = fmaf(m11, m22, -(m12 * m12));
    p.all = !(m11 > 0.0f) || !(det > 1e-4f * (m11 * m22));
    p.u0 = (r[0] - bx) * inv_step;
    p.dy0 = by - r[1];
    p.s = (m12 * inv22) * inv_step;
    p.a = ((det * inv22) * inv22) * (inv_step * inv_step);
    p.m22 = m22;
    return p;
}

// w(y)^2 in columns^2 for row offset row_dy (negative: the row misses).
FH_HD float fh_mask_row_w2(const FhMaskPrep& p, float row_dy) {
    return p.a * fmaf(-row_dy, row_dy, p.m22);
}

// clamp(x, lo, hi) for lo <= hi and x not NaN: v_med3_f32 on the GPU, the
// same value as max(lo, min(x, hi)) on the host.
FH_HD float fh_med3(float x, float lo, float hi) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_fmed3f(x, lo, hi);
#else
    return fh_fmax(lo, fh_fmin(x, hi));
#endif
}

// The column bits (at most 8 columns) of the row at y offset row_dy from the
// centre, given the row's half width in columns already widened by the
// margin (sqrt(fh_mask_row_w2) + kMaskMarginPx), or -1 when fh_mask_row_w2
// is negative (the row misses: the columns then form an empty range).
FH_HD uint32_t fh_mask_row_bits(const FhMaskPrep& p, float row_dy, float w, int cols) {
    const float uc = fmaf(p.s, row_dy, p.u0);
    const float fc = (float)cols;
    // columns [b, e): b = ceil(uc - w), e = floor(uc + w) + 1, both in [0, cols]
    const float b = fh_med3(ceilf(uc - w), 0.0f, fc);
    const float e = fh_med3(floorf(uc + w) + 1.0f, 0.0f, fc);
    const uint32_t ib = (uint32_t)b, n = (uint32_t)fh_fmax(e - b, 0.0f);
#if defined(__HIP_DEVICE_COMPILE__)
    // n, ib <= cols <= 8: v_bfm_b32 gives ((1 << n) - 1) << ib in one instruction
    uint32_t m;
    __asm__("v_bfm_b32 %0, %1, %2" : "=v"(m) : "v"(n), "v"(ib));
    return m;
#else
    return ((1u << n) - 1u) << ib;
#endif
}
