// K-buffer hybrid-transparency blend (rasterization/include/kernels/inference.cuh:138-286).
//
// One 64-lane workgroup (a single wave) per active 8x8 small tile, any tile
// type in one launch: FOVEA and BLENDED tiles are full resolution, PERIPHERY
// tiles cover their 16x16 large tile at stride 2 (inference.cuh:166-169).
// Gaussians of the tile's range are staged through LDS in batches of 64; per
// pixel the K nearest fragments with alpha >= 0.05 are kept sorted in
// registers with 8-bit quantised colour, everything else is accumulated
// order-independently in the tail; resolve, optional background, saturate.
// Extension: QUARTER tiles cover their 16x16 large tile at stride 4 (4x4
// samples), four tiles per wave: 16-lane groups, each walking its own tile's
// range in batches of 16 through its own quarter of the LDS batch.
// Per LDS batch each lane first marks the staged Gaussians that can reach its
// pixel -- one bit of each staged Gaussian's sample mask, which its staging
// lane computed from the support ellipse row by row -- then evaluates only
// those, in order (two passes, same arithmetic).
// Workgroups take their tiles from a work order with the longest instance
// lists first (blend_order_count_kernel, blend_order_place_kernel).
// Work tiles: with W = 2 or 4 waves per workgroup (blend_wide_kernel) the
// sub-tiles of a FOVEA or BLENDED large tile, whose instance lists are the
// same, are rendered as one 16x8 or 16x16 work tile: its waves stage the list
// once, 64 W Gaussians a round, one lane per Gaussian, each 16-column row
// interval of the ellipse serving both sub-tiles of its row; each wave then
// walks its own sub-tile's masks. PERIPHERY tiles go W to a workgroup and
// QUARTER tiles 4 W, one wave each as in the single-wave kernel. A pixel sees
// the same Gaussians in the same order with every work tile.
#include "fhtgs_device.h"
#include "fhtgs_cull.h"

namespace fhtgs {

// sh_utils.cuh:9-30 (background SH)
constexpr float B_C0 = 0.28209479177387814f;
constexpr float B_C1 = 0.48860251190291987f;
constexpr float B_C2a = 1.0925484305920792f;
constexpr float B_C2b = 0.94617469575755997f;
constexpr float B_C2c = 0.31539156525251999f;
constexpr float B_C2d = 0.54627421529603959f;
constexpr float B_C3a = 0.59004358992664352f;
constexpr float B_C3b = 1.7701307697799304f;
constexpr float B_C3c = 2.8906114426405538f;
constexpr float B_C3d = 0.45704579946446572f;
constexpr float B_C3e = 2.2852289973223288f;
constexpr float B_C3f = 1.865881662950577f;
constexpr float B_C3g = 1.1195289977703462f;
constexpr float B_C3h = 1.4453057213202769f;

__device__ __forceinline__ v3 pixel_dir(const ViewParams& vp, float px, float py) {
    const v4 pc{px, py, 1.0f, 1.0f};
    const v3 t{dot4(load_row(vp.VPR_inv, 0), pc), dot4(load_row(vp.VPR_inv, 1), pc),
               dot4(load_row(vp.VPR_inv, 2), pc)};
    return normalize3(t);
}

// background_model.cuh:9-48
__device__ v3 sh_background(const ViewParams& vp, float px, float py) {
    const v3 t = pixel_dir(vp, px, py);
    if (t.y > 0.1f) return v3{0.0f, 0.0f, 0.0f};
    const v3 n = normalize3(t);
    const float x = n.x, y = n.y, z = n.z;
    const float xx = x * x, yy = y * y, zz = z * z;
    const float xy = x * y, xz = x * z, yz = y * z;
    const float* k = vp.bg_data;
    const float f[15] = {
        -(B_C1 * y), B_C1 * z, -(B_C1 * x),
        B_C2a * xy, -(B_C2a * yz), fmaf(B_C2b, zz, -B_C2c), -(B_C2a * xz), B_C2d * (xx - yy),
        y * fmaf(B_C3a, yy, -(B_C3b * xx)), (B_C3c * xy) * z, y * fmaf(-B_C3e, zz, B_C3d),
        z * fmaf(B_C3f, zz, -B_C3g), x * fmaf(-B_C3e, zz, B_C3d), (B_C3h * z) * (xx - yy),
        x * fmaf(B_C3b, yy, -(B_C3a * xx))};
    float r0 = fmaf(B_C0, k[0], 0.5f), r1 = fmaf(B_C0, k[1], 0.5f), r2 = fmaf(B_C0, k[2], 0.5f);
#pragma unroll
    for (int j = 0; j < 15; ++j) {
        r0 = fmaf(f[j], k[(j + 1) * 3 + 0], r0);
        r1 = fmaf(f[j], k[(j + 1) * 3 + 1], r1);
        r2 = fmaf(f[j], k[(j + 1) * 3 + 2], r2);
    }
    v3 col{fh_saturate(fmaf(fh_tanhf(r0), 0.5f, 0.5f)), fh_saturate(fmaf(fh_tanhf(r1), 0.5f, 0.5f)),
           fh_saturate(fmaf(fh_tanhf(r2), 0.5f, 0.5f))};
    if (t.y > 0.0f) {
        const float s = 1.0f - t.y / 0.1f;
        col = v3{s * col.x, s * col.y, s * col.z};
    }
    return col;
}

// background_model.cuh:52-117
__device__ v3 tex_background(const ViewParams& vp, float px, float py, bool activate) {
    const int W = vp.bg_w, H = vp.bg_h;
    const v3 t = pixel_dir(vp, px, py);
    if (t.y > 0.1f) return v3{0.0f, 0.0f, 0.0f};
    const float pi = 0x1.921fb6p+1f;
    const float ex = fmaf(fh_atan2f(t.x, t.z) / pi, 0.5f, 0.5f);
    const float ey = fmaf(fh_asinf(t.y) / (0.5f * pi), 0.5f, 0.5f);
    const float tx = fmaf(ex, (float)W, -0.5f);
    const float ty = fmaf(ey, (float)H, -0.5f);
    const int x0 = fh_f2i_rd(tx), y0 = fh_f2i_rd(ty);
    const int x1 = x0 + 1, y1 = y0 + 1;
    const float fx = tx - (float)x0, fy = ty - (float)y0;
    const int x0w = x0 < 0 ? x0 + W : x0;
    const int x1w = x1 >= W ? x1 - W : x1;
    const int y0c = max(0, min(H - 1, y0));
    const int y1c = max(0, min(H - 1, y1));
    const int64_t i00 = x0w + (int64_t)y0c * W, i10 = x1w + (int64_t)y0c * W;
    const int64_t i01 = x0w + (int64_t)y1c * W, i11 = x1w + (int64_t)y1c * W;
    const int64_t nt = (int64_t)W * H;
    float c[3];
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
        const float* d = vp.bg_data + ch * nt;
        const float c0 = fmaf(d[i00], 1.0f - fx, d[i10] * fx);
        const float c1 = fmaf(d[i01], 1.0f - fx, d[i11] * fx);
        float val = fmaf(c0, 1.0f - fy, c1 * fy);
        if (activate) val = 1.0f / (1.0f + fh_expf(-val));
        c[ch] = val;
    }
    v3 col{c[0], c[1], c[2]};
    if (t.y > 0.0f) {
        const float s = 1.0f - t.y / 0.1f;
        col = v3{s * col.x, s * col.y, s * col.z};
    }
    return col;
}

// FragmentInfo::pack_channel (inference.cuh:20-22)
__device__ __forceinline__ uint32_t pack_channel(float x) {
    return (uint32_t)fmaf(fminf(fmaxf(x, 0.0f), 1.0f), 255.0f, 0.5f);
}

struct BlendArgs {
    const uint32_t* frame_info;
    const uint32_t* order;  // work units, longest first (blend_order_place_kernel)
    const uint32_t* tile_map_part;
    const uint2* ranges;
    const uint32_t* vals;
    const float4* prim;  // 4 float4 per primitive: VPMT1, VPMT2, VPMT4, rgba
    const float4* ell;  // per primitive, 32 B: support-ellipse record (fh_ellipse_record):
                        // (cx, cy, m11, m22), (m12, 0, 0, 0)
    uint32_t xcd_run;
    uint32_t n_small;   // bound of every partition range
};

// The dispatcher hands workgroups to the 8 XCDs round-robin by index, and
// each XCD has its own L2. Within every group of 8*run workgroups, workgroup
// b becomes tile (b % 8) * run + b / 8 of the group, so each XCD walks runs
// of `run` consecutive tiles (neighbours share most of their Gaussians). A
// trailing partial group keeps its order; the map is a bijection on [0, n).
constexpr uint32_t kNumXcd = 8;
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t n, uint32_t run) {
    if (run <= 1u) return b;
    const uint32_t span = kNumXcd * run;
    const uint32_t grp = b / span;
    if ((uint64_t)(grp + 1u) * span > n) return b;
    const uint32_t w = b - grp * span;
    return grp * span + (w % kNumXcd) * run + w / kNumXcd;
}

// One pixel's blend state: the K nearest core fragments (depth, 8-bit
// quantised rgba) sorted in registers, and the order-independent tail.
template <int K>
struct PixelState {
    float cd[K];
    uint32_t cc[K];
    float T_tail, t0, t1, t2, t3;
};

template <int K>
__device__ __forceinline__ void pixel_init(PixelState<K>& ps) {
#pragma unroll
    for (int c = 0; c < K; ++c) { ps.cd[c] = 3.402823466e+38f; ps.cc[c] = 0u; }
    ps.T_tail = 1.0f;
    ps.t0 = ps.t1 = ps.t2 = ps.t3 = 0.0f;
}

// The reference's evaluation (inference.cuh:200-238) of one staged Gaussian
// at pixel (pxf, pyf): `off` is its byte offset in the LDS arrays of VPMT1,
// VPMT2, VPMT4 and rgba. Cut-off, correctly rounded reciprocal, exponential,
// alpha test, the core insertion (strict-< swap chain) or the tail.
// The work tiles' shared staging (blend_shared) uses these functions;
// blend_wave, which the single-wave kernel runs, keeps the same operations
// as lambdas over its own locals: built on these functions its register
// allocation grew from 92 to 96 VGPRs with spills at the five-waves-per-SIMD
// budget, and the blend took 2 % longer.
template <int K>
__device__ __forceinline__ void pixel_eval(PixelState<K>& ps, const float4* sV1, const float4* sV2,
                                           const float4* sV4, const float4* sC, uint32_t off, float pxf,
                                           float pyf) {
    auto at = [&](const float4* arr) -> float4 {
        return *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(arr) + off);
    };
    const float4 W1 = at(sV1), W2 = at(sV2), W4 = at(sV4);
    const float Px0 = fmaf(-W4.x, pxf, W1.x), Px1 = fmaf(-W4.y, pxf, W1.y);
    const float Px2 = fmaf(-W4.z, pxf, W1.z), Px3 = fmaf(-W4.w, pxf, W1.w);
    const float Py0 = fmaf(-W4.x, pyf, W2.x), Py1 = fmaf(-W4.y, pyf, W2.y);
    const float Py2 = fmaf(-W4.z, pyf, W2.z), Py3 = fmaf(-W4.w, pyf, W2.w);
    const v3 m{fmaf(Px3, Py0, -(Px0 * Py3)), fmaf(Px3, Py1, -(Px1 * Py3)), fmaf(Px3, Py2, -(Px2 * Py3))};
    const v3 d = cross3(v3{Px0, Px1, Px2}, v3{Py0, Py1, Py2});
    const float num = dot3(m, m);
    const float den = dot3(d, d);
    if (num > kMaxCutoffSq * den) return;
    // num <= 11.08 den here, so the exponent lies in [-6, 0] (or is NaN)
    const float rcp = rcp_rn(den);
    const float G = fh_expf_m64_gfx((-0.5f * num) * rcp);
    const float4 Cj = at(sC);
    const float alpha = fminf(Cj.w * G, kMaxFragmentAlpha);
    if (alpha < kMinAlpha) return;
    float p0 = Cj.x * alpha, p1 = Cj.y * alpha, p2 = Cj.z * alpha, p3 = alpha;
    // the core test of inference.cuh:225 with its alpha half first: the
    // depth only matters for fragments that may enter the core
    if (alpha >= kMinAlphaCore) {
        const v3 e0 = cross3(d, m);
        const float depth = fmaf(W4.z, e0.z * rcp, fmaf(W4.y, e0.y * rcp, W4.x * (e0.x * rcp))) + W4.w;
        if (depth < ps.cd[K - 1]) {
            float fd = depth;
            uint32_t fc = pack_channel(p0) | (pack_channel(p1) << 8) | (pack_channel(p2) << 16) |
                          (pack_channel(p3) << 24);
            // insertion into the sorted core (swap chain of inference.cuh:227-232), branch-free
#pragma unroll
            for (int c = 0; c < K; ++c) {
                const bool lt = fd < ps.cd[c];
                const float nd = lt ? ps.cd[c] : fd;  // carried on
                const uint32_t nc = lt ? ps.cc[c] : fc;
                ps.cd[c] = lt ? fd : ps.cd[c];
                ps.cc[c] = lt ? fc : ps.cc[c];
                fd = nd;
                fc = nc;
            }
            const float s = 1.0f / 255.0f;
            p0 = (float)(fc & 0xffu) * s;
            p1 = (float)((fc >> 8) & 0xffu) * s;
            p2 = (float)((fc >> 16) & 0xffu) * s;
            p3 = (float)((fc >> 24) & 0xffu) * s;
        }
    }
    const float tt = 1.0f - p3;
    ps.t0 = ps.t0 + p0; ps.t1 = ps.t1 + p1; ps.t2 = ps.t2 + p2; ps.t3 = ps.t3 + p3;
    ps.T_tail = ps.T_tail * tt;
}

// Both passes over staged entries [e0, e0 + n), n <= 64, of one batch.
// Pass 1: which of them can reach this pixel: this lane's bit (m_word,
// m_shift) of each entry's sample mask (two words per entry at msk), a
// superset of the pairs the reference's cutoff and alpha tests keep. Bit jj
// of ra is entry e0 + jj, bit jj - 32 of rb entry e0 + jj (jj >= 32).
// Pass 2: the reference's evaluation of those, each lane walking its own bits
// in ascending order, so every pixel sees the same Gaussians in the same
// order as a walk over the whole batch, and the wave runs as many steps as
// its busiest lane needs instead of one per staged Gaussian.
template <int K>
__device__ __forceinline__ void blend_batch(PixelState<K>& ps, const uint32_t* msk, uint32_t e0, int n,
                                            uint32_t m_word, uint32_t m_shift, const float4* sV1,
                                            const float4* sV2, const float4* sV4, const float4* sC, float pxf,
                                            float pyf) {
    auto may_reach = [&](uint32_t e) -> uint32_t { return __builtin_amdgcn_ubfe(msk[2u * e + m_word], m_shift, 1u); };
    const int na = n < 32 ? n : 32;
    uint32_t ra = 0u, rb = 0u;
#pragma unroll 4
    for (int jj = 0; jj < na; ++jj) ra |= may_reach(e0 + (uint32_t)jj) << jj;
#pragma unroll 4
    for (int jj = 32; jj < n; ++jj) rb |= may_reach(e0 + (uint32_t)jj) << (jj - 32);
    // the batch's entries as byte offsets from each array
    const uint32_t ebyte = e0 * (uint32_t)sizeof(float4);
    // lowest set bit first, i.e. in list order
    if (n <= 32) {
        // one word (the usual case)
        uint32_t reach = ra;
        while (reach != 0u) {
            const uint32_t jj = (uint32_t)__builtin_ctz(reach);
            reach &= reach - 1u;
            pixel_eval<K>(ps, sV1, sV2, sV4, sC, ebyte + (jj << 4), pxf, pyf);
        }
    } else {
        uint64_t reach = (uint64_t)ra | ((uint64_t)rb << 32);
        while (reach != 0ull) {
            const uint32_t jj = (uint32_t)__builtin_ctzll(reach);
            reach &= reach - 1ull;
            pixel_eval<K>(ps, sV1, sV2, sV4, sC, ebyte + (jj << 4), pxf, pyf);
        }
    }
}

// Resolve (inference.cuh:240-286): the core front to back, the tail's
// weighted average behind it, the background, saturate; FOVEA tiles write the
// final image, the others the blend target.
template <int K>
__device__ __forceinline__ void pixel_resolve_store(const PixelState<K>& ps, const ViewParams& vp, int type,
                                                    uint32_t px, uint32_t py, float pxf, float pyf) {
    float r = 0.0f, gch = 0.0f, bch = 0.0f;
    float Tc = 1.0f;
    const float s255 = 1.0f / 255.0f;
    // the loop condition of inference.cuh:246; once it fails Tc no longer changes
#pragma unroll
    for (int c = 0; c < K; ++c) {
        const bool on = Tc >= kTransmittanceThreshold;
        const uint32_t col = ps.cc[c];
        const float q0 = (float)(col & 0xffu) * s255, q1 = (float)((col >> 8) & 0xffu) * s255;
        const float q2 = (float)((col >> 16) & 0xffu) * s255, q3 = (float)((col >> 24) & 0xffu) * s255;
        r = on ? fmaf(Tc, q0, r) : r;
        gch = on ? fmaf(Tc, q1, gch) : gch;
        bch = on ? fmaf(Tc, q2, bch) : bch;
        Tc = on ? Tc * (1.0f - q3) : Tc;
    }
    if (Tc >= kTransmittanceThreshold && ps.t3 >= kMinAlpha) {
        const float wt = Tc * (1.0f - ps.T_tail);
        const float s = wt * (1.0f / ps.t3);
        r = fmaf(s, ps.t0, r);
        gch = fmaf(s, ps.t1, gch);
        bch = fmaf(s, ps.t2, bch);
    }
    if (vp.bg_type != 0) {
        const float wb = Tc * ps.T_tail;
        if (wb >= kTransmittanceThreshold) {
            const v3 bg = vp.bg_type == 1 ? sh_background(vp, pxf, pyf)
                                          : tex_background(vp, pxf, pyf, vp.bg_type == 3);
            r = fmaf(wb, bg.x, r);
            gch = fmaf(wb, bg.y, gch);
            bch = fmaf(wb, bg.z, bch);
        }
    }
    float* out = type == kTileFovea ? vp.image_final : vp.image_blend;
    const int64_t idx = (int64_t)vp.width * py + px;
    if (vp.to_chw) {
        const int64_t np = (int64_t)vp.width * vp.height;
        out[idx] = fh_saturate(r);
        out[np + idx] = fh_saturate(gch);
        out[2 * np + idx] = fh_saturate(bch);
    } else {
        out[3 * idx] = fh_saturate(r);
        out[3 * idx + 1] = fh_saturate(gch);
        out[3 * idx + 2] = fh_saturate(bch);
    }
}

// Orders a wave's LDS accesses on both sides (a wave's LDS operations are
// performed in order), for waves that share no LDS with the rest of their
// workgroup.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One wave renders one FOVEA, PERIPHERY or BLENDED tile, or four QUARTER
// tiles (16 lanes each): j is this lane's partitioned tile, `have` false for
// the lanes of a QUARTER group without one. It stages 64 Gaussians a round
// through its own LDS arrays (sV1, sV2, sV4, sC: 64 float4 each; sM: 128
// words). kWaveSync: the workgroup has other waves, which share nothing with
// this one, so the staging is ordered within the wave only.
template <int K, bool kWaveSync>
__device__ __forceinline__ void blend_wave(const ViewParams& vp, const BlendArgs& a, int type, uint32_t j,
                                           bool have, uint32_t lane, float4* sV1, float4* sV2, float4* sV4,
                                           float4* sC, uint32_t* sM) {
    auto sync = [&]() {
        if constexpr (kWaveSync) wave_sync();
        else __syncthreads();
    };
    // lanes form groups of `gsz` with a shared tile and range: the whole wave,
    // or 16 lanes per QUARTER tile (the last wave may have empty groups)
    const bool quarter = type == kTileQuarter;
    const uint32_t gsz = quarter ? 16u : 64u;
    const uint32_t gbase = quarter ? (lane & ~15u) : 0u;
    const uint32_t li = lane - gbase;
    const uint64_t gmask = quarter ? (0xffffull << gbase) : ~0ull;
    const uint32_t tg = have ? a.tile_map_part[j] : 0u;
    const uint32_t lt = tg / kNumSub, st = tg % kNumSub;
    const uint32_t ltx = lt % (uint32_t)vp.gw, lty = lt / (uint32_t)vp.gw;
    const uint32_t step = quarter ? (uint32_t)kQuarterStride : (type == kTilePeriphery ? kStride : 1u);
    const uint32_t cols = quarter ? 4u : 8u;  // samples per row
    const uint32_t bx = ltx * kTileL + (st % kStride) * kTileS;
    const uint32_t by = lty * kTileL + (st / kStride) * kTileS;
    const uint32_t px = bx + (li % cols) * step;
    const uint32_t py = by + (li / cols) * step;
    const bool inside = have && px < (uint32_t)vp.width && py < (uint32_t)vp.height;
    const float pxf = (float)px, pyf = (float)py;
    const float tx0 = (float)bx, ty0 = (float)by;  // the group's first sample
    const float stepf = (float)step, inv_step = quarter ? 0.25f : (step == 2u ? 0.5f : 1.0f);
    // this lane's bit in its group's sample masks (first pass)
    const uint32_t m_word = li >> 5, m_shift = li & 31u;
    const uint64_t lt_mask = lanemask_lt() & gmask;

    float cd[K];
    uint32_t cc[K];
#pragma unroll
    for (int c = 0; c < K; ++c) { cd[c] = 3.402823466e+38f; cc[c] = 0u; }
    float T_tail = 1.0f;
    float t0 = 0.0f, t1 = 0.0f, t2 = 0.0f, t3 = 0.0f;

    const uint2 range = have ? a.ranges[tg] : make_uint2(0u, 0u);
    for (uint32_t b0 = range.x; __ballot(b0 < range.y) != 0ull; b0 += gsz) {
        // each lane loads one instance of its group's next gsz and computes
        // its sample mask from the support ellipse (the preprocess's record):
        // which of the group's samples the ellipse can reach, row by row
        // (fh_mask_row_bits; every sample for a record without a usable
        // ellipse). A Gaussian whose mask holds no sample reaches none and is
        // dropped (the tile cull); the others are compacted in order into the
        // group's part of the LDS batch with their masks.
        const uint32_t idx = b0 + li;
        float4 V1 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), V2 = V1, V4 = V1, C = V1;
        uint32_t m0 = 0u, m1 = 0u;
        if (idx < range.y) {
            const uint32_t p = a.vals[idx];
            const float4 E = a.ell[2 * (size_t)p];
            const float er[5] = {E.x, E.y, E.z, E.w, a.ell[2 * (size_t)p + 1].x};
            const float4* rec = a.prim + 4 * (size_t)p;
            V1 = rec[0];
            V2 = rec[1];
            V4 = rec[2];
            C = rec[3];
            const FhMaskPrep mp = fh_mask_prep(er, rcp_rn(er[3]), tx0, ty0, inv_step);
            if (mp.all) {
                m0 = quarter ? 0xffffu : ~0u;
                m1 = quarter ? 0u : ~0u;
            } else {
#pragma unroll
                for (uint32_t r = 0; r < 8u; ++r) {
                    if (r >= cols) continue;
                    const float dy = fmaf((float)r, stepf, mp.dy0);
                    const float w2 = fh_mask_row_w2(mp, dy);
                    const float w = w2 < 0.0f ? -1.0f : __builtin_amdgcn_sqrtf(w2) + kMaskMarginPx;
                    const uint32_t bits = fh_mask_row_bits(mp, dy, w, (int)cols);
                    const uint32_t sh = r * cols;
                    if (sh < 32u) m0 |= bits << sh;
                    else m1 |= bits << (sh - 32u);
                }
            }
        }
        const bool keep = (m0 | m1) != 0u;
        const uint64_t km = __ballot(keep);
        if (km == 0ull) continue;
        const int count = (int)__popcll(km & gmask);
        // the first pass runs the same number of steps in every lane: the
        // largest group count of the wave (QUARTER groups differ)
        int nmax = (int)__popcll(km);
        if (quarter) {
            const int c0 = (int)__popcll(km & 0xffffull), c1 = (int)__popcll((km >> 16) & 0xffffull);
            const int c2 = (int)__popcll((km >> 32) & 0xffffull), c3 = (int)__popcll(km >> 48);
            nmax = max(max(c0, c1), max(c2, c3));
        }
        sync();
        const uint32_t pos = gbase + (uint32_t)__popcll(km & lt_mask);
        if (keep) {
            sV1[pos] = V1;
            sV2[pos] = V2;
            sV4[pos] = V4;
            sC[pos] = C;
            sM[2u * pos] = m0;
            sM[2u * pos + 1u] = m1;
        }
        if ((int)li >= count) {
            // a slot past the group's Gaussians (any lane there, staging or
            // not, since its Gaussian went to a lower slot): a mask with no
            // sample, for the steps of groups with fewer Gaussians than the
            // wave's largest group
            sM[2u * (gbase + li)] = 0u;
            sM[2u * (gbase + li) + 1u] = 0u;
        }
        sync();
        if (!inside) continue;
        // Pass 1: which staged Gaussians can reach this pixel: this lane's bit
        // of each staged Gaussian's sample mask, a superset of the pairs the
        // reference's cutoff and alpha tests keep. Bit jj of ra is staged
        // Gaussian jj, bit jj - 32 of rb staged Gaussian jj (jj >= 32).
        auto may_reach = [&](uint32_t e) -> uint32_t { return __builtin_amdgcn_ubfe(sM[2u * e + m_word], m_shift, 1u); };
        const int na = nmax < 32 ? nmax : 32;
        uint32_t ra = 0u, rb = 0u;
#pragma unroll 4
        for (int jj = 0; jj < na; ++jj) ra |= may_reach(gbase + (uint32_t)jj) << jj;
#pragma unroll 4
        for (int jj = 32; jj < nmax; ++jj) rb |= may_reach(gbase + (uint32_t)jj) << (jj - 32);
        // Pass 2: the reference's evaluation (inference.cuh:200-238) of those
        // Gaussians, each lane walking its own bits in ascending order, so
        // every pixel sees the same Gaussians in the same order as a walk
        // over the whole batch, and the wave runs as many steps as its
        // busiest lane needs instead of one per staged Gaussian.
        // the group's part of the batch, as byte offsets from each array
        const uint32_t gbyte = gbase * (uint32_t)sizeof(float4);
        auto slot = [&](const float4* arr, uint32_t jj) -> float4 {
            return *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(arr) + (gbyte + (jj << 4)));
        };
        auto evaluate = [&](uint32_t jj) {
            const float4 W1 = slot(sV1, jj), W2 = slot(sV2, jj), W4 = slot(sV4, jj);
            const float Px0 = fmaf(-W4.x, pxf, W1.x), Px1 = fmaf(-W4.y, pxf, W1.y);
            const float Px2 = fmaf(-W4.z, pxf, W1.z), Px3 = fmaf(-W4.w, pxf, W1.w);
            const float Py0 = fmaf(-W4.x, pyf, W2.x), Py1 = fmaf(-W4.y, pyf, W2.y);
            const float Py2 = fmaf(-W4.z, pyf, W2.z), Py3 = fmaf(-W4.w, pyf, W2.w);
            const v3 m{fmaf(Px3, Py0, -(Px0 * Py3)), fmaf(Px3, Py1, -(Px1 * Py3)),
                       fmaf(Px3, Py2, -(Px2 * Py3))};
            const v3 d = cross3(v3{Px0, Px1, Px2}, v3{Py0, Py1, Py2});
            const float num = dot3(m, m);
            const float den = dot3(d, d);
            if (num > kMaxCutoffSq * den) return;
            // num <= 11.08 den here, so the exponent lies in [-6, 0] (or is NaN)
            const float rcp = rcp_rn(den);
            const float G = fh_expf_m64_gfx((-0.5f * num) * rcp);
            const float4 Cj = slot(sC, jj);
            const float alpha = fminf(Cj.w * G, kMaxFragmentAlpha);
            if (alpha < kMinAlpha) return;
            float p0 = Cj.x * alpha, p1 = Cj.y * alpha, p2 = Cj.z * alpha, p3 = alpha;
            // the core test of inference.cuh:225 with its alpha half first: the
            // depth only matters for fragments that may enter the core
            if (alpha >= kMinAlphaCore) {
                const v3 e0 = cross3(d, m);
                const float depth = fmaf(W4.z, e0.z * rcp, fmaf(W4.y, e0.y * rcp, W4.x * (e0.x * rcp))) + W4.w;
                if (depth < cd[K - 1]) {
                    float fd = depth;
                    uint32_t fc = pack_channel(p0) | (pack_channel(p1) << 8) | (pack_channel(p2) << 16) |
                                  (pack_channel(p3) << 24);
                    // insertion into the sorted core (swap chain of inference.cuh:227-232), branch-free
#pragma unroll
                    for (int c = 0; c < K; ++c) {
                        const bool lt = fd < cd[c];
                        const float nd = lt ? cd[c] : fd;  // carried on
                        const uint32_t nc = lt ? cc[c] : fc;
                        cd[c] = lt ? fd : cd[c];
                        cc[c] = lt ? fc : cc[c];
                        fd = nd;
                        fc = nc;
                    }
                    const float s = 1.0f / 255.0f;
                    p0 = (float)(fc & 0xffu) * s;
                    p1 = (float)((fc >> 8) & 0xffu) * s;
                    p2 = (float)((fc >> 16) & 0xffu) * s;
                    p3 = (float)((fc >> 24) & 0xffu) * s;
                }
            }
            const float tt = 1.0f - p3;
            t0 = t0 + p0; t1 = t1 + p1; t2 = t2 + p2; t3 = t3 + p3;
            T_tail = T_tail * tt;
        };
        // lowest set bit first, i.e. in list order
        if (nmax <= 32) {
            // one word (the usual case)
            uint32_t reach = ra;
            while (reach != 0u) {
                const uint32_t jj = (uint32_t)__builtin_ctz(reach);
                reach &= reach - 1u;
                evaluate(jj);
            }
        } else {
            uint64_t reach = (uint64_t)ra | ((uint64_t)rb << 32);
            while (reach != 0ull) {
                const uint32_t jj = (uint32_t)__builtin_ctzll(reach);
                reach &= reach - 1ull;
                evaluate(jj);
            }
        }
    }
    if (!inside) return;
    float r = 0.0f, gch = 0.0f, bch = 0.0f;
    float Tc = 1.0f;
    const float s255 = 1.0f / 255.0f;
    // the loop condition of inference.cuh:246; once it fails Tc no longer changes
#pragma unroll
    for (int c = 0; c < K; ++c) {
        const bool on = Tc >= kTransmittanceThreshold;
        const uint32_t col = cc[c];
        const float q0 = (float)(col & 0xffu) * s255, q1 = (float)((col >> 8) & 0xffu) * s255;
        const float q2 = (float)((col >> 16) & 0xffu) * s255, q3 = (float)((col >> 24) & 0xffu) * s255;
        r = on ? fmaf(Tc, q0, r) : r;
        gch = on ? fmaf(Tc, q1, gch) : gch;
        bch = on ? fmaf(Tc, q2, bch) : bch;
        Tc = on ? Tc * (1.0f - q3) : Tc;
    }
    if (Tc >= kTransmittanceThreshold && t3 >= kMinAlpha) {
        const float wt = Tc * (1.0f - T_tail);
        const float s = wt * (1.0f / t3);
        r = fmaf(s, t0, r);
        gch = fmaf(s, t1, gch);
        bch = fmaf(s, t2, bch);
    }
    if (vp.bg_type != 0) {
        const float wb = Tc * T_tail;
        if (wb >= kTransmittanceThreshold) {
            const v3 bg = vp.bg_type == 1 ? sh_background(vp, pxf, pyf)
                                          : tex_background(vp, pxf, pyf, vp.bg_type == 3);
            r = fmaf(wb, bg.x, r);
            gch = fmaf(wb, bg.y, gch);
            bch = fmaf(wb, bg.z, bch);
        }
    }
    float* out = type == kTileFovea ? vp.image_final : vp.image_blend;
    const int64_t idx = (int64_t)vp.width * py + px;
    if (vp.to_chw) {
        const int64_t np = (int64_t)vp.width * vp.height;
        out[idx] = fh_saturate(r);
        out[np + idx] = fh_saturate(gch);
        out[2 * np + idx] = fh_saturate(bch);
    } else {
        out[3 * idx] = fh_saturate(r);
        out[3 * idx + 1] = fh_saturate(gch);
        out[3 * idx + 2] = fh_saturate(bch);
    }
}

// 8x8 work tiles: one single-wave workgroup per unit of the work order (a
// FOVEA, PERIPHERY or BLENDED tile, or four QUARTER tiles).
template <int K>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(K <= 8 ? 5 : 1))) void blend_kernel(const ViewParams* vpp, BlendArgs a) {
    const ViewParams vp = load_view(vpp);
    __shared__ float4 sV1[64], sV2[64], sV4[64], sC[64];
    // first-pass data per staged Gaussian: its sample mask as two words
    // (bit row * cols + col, fh_mask_row_bits), or every sample for a record
    // without a usable ellipse
    __shared__ uint32_t sM[128];
    const uint32_t* fi = a.frame_info;
    const uint32_t g0 = xcd_remap(blockIdx.x, gridDim.x, a.xcd_run);
    const uint32_t nq = fi[7] - fi[6];
    if (g0 >= fi[FI_N_ACTIVE] - nq + (nq + 3u) / 4u) return;
    const uint32_t g = a.order[g0];  // this workgroup's unit
    const uint32_t lane = threadIdx.x;
    int type;
    uint32_t j = g;  // this lane's partitioned tile
    if (g >= fi[0] && g < fi[1]) type = kTileFovea;
    else if (g >= fi[2] && g < fi[3]) type = kTilePeriphery;
    else if (g >= fi[4] && g < fi[5]) type = kTileBlended;
    else if (g >= fi[6] && g < fi[6] + (fi[7] - fi[6] + 3u) / 4u) {
        type = kTileQuarter;
        j = fi[6] + (g - fi[6]) * 4u + (lane >> 4);
    } else return;
    blend_wave<K, false>(vp, a, type, j, type != kTileQuarter || j < fi[7], lane, sV1, sV2, sV4, sC, sM);
}

// Work units with W waves per workgroup (W = 2: 16x8 work tiles, W = 4:
// 16x16), in partition order: FOVEA work tiles (W consecutive partitioned
// sub-tiles of one large tile: a FOVEA or BLENDED large tile contributes its
// four sub-tiles in order), PERIPHERY tiles W at a time, BLENDED work tiles,
// QUARTER tiles 4 W at a time. Every range is clamped to n_small.
struct WideUnits {
    uint32_t bP, bB, bQ, end;  // first unit of each type after FOVEA; units in total
    uint32_t f0, p0, b0, q0;   // first partitioned entry of each type
    uint32_t pe, qe;           // end of the PERIPHERY and QUARTER entries
};

__device__ __forceinline__ WideUnits wide_units(const uint32_t* fi, uint32_t n_small, uint32_t W) {
    WideUnits U;
    uint32_t e[4];
    uint32_t* s[4] = {&U.f0, &U.p0, &U.b0, &U.q0};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        e[k] = min(fi[FI_PART + 2 * k + 1], n_small);
        *s[k] = min(fi[FI_PART + 2 * k], e[k]);
    }
    U.pe = e[1];
    U.qe = e[3];
    U.bP = (e[0] - U.f0) / W;
    U.bB = U.bP + (e[1] - U.p0 + W - 1u) / W;
    U.bQ = U.bB + (e[2] - U.b0) / W;
    U.end = U.bQ + (e[3] - U.q0 + 4u * W - 1u) / (4u * W);
    return U;
}

struct WideUnit {
    int type;
    uint32_t first, count;  // partitioned entries [first, first + count)
};

__device__ __forceinline__ WideUnit wide_unit(const WideUnits& U, uint32_t u, uint32_t W) {
    WideUnit r;
    if (u < U.bP) {
        r.type = kTileFovea;
        r.first = U.f0 + W * u;
        r.count = W;
    } else if (u < U.bB) {
        r.type = kTilePeriphery;
        r.first = U.p0 + W * (u - U.bP);
        r.count = min(W, U.pe - r.first);
    } else if (u < U.bQ) {
        r.type = kTileBlended;
        r.first = U.b0 + W * (u - U.bB);
        r.count = W;
    } else {
        r.type = kTileQuarter;
        r.first = U.q0 + 4u * W * (u - U.bQ);
        r.count = min(4u * W, U.qe - r.first);
    }
    return r;
}

// LDS of blend_wide_kernel, in float4: per wave, 64 staged Gaussians' VPMT1,
// VPMT2, VPMT4, rgba and 128 mask words (single-wave units); or, for a work
// tile, 64 W staged Gaussians, W sub-tile masks of two words each per
// Gaussian and the round's per-wave keep ballots (two rounds apart).
constexpr uint32_t kWaveLdsF4 = 4u * 64u + 128u / 4u;
template <int W>
constexpr uint32_t wide_lds_f4() {
    const uint32_t shared = 4u * 64u * W + (2u * W * 64u * W) / 4u + (2u * W * 2u) / 4u + 1u;
    const uint32_t waves = kWaveLdsF4 * W;
    return shared > waves ? shared : waves;
}

// One 16x8 (W = 2) or 16x16 (W = 4) work tile of a FOVEA or BLENDED large
// tile, entries j0 .. j0 + W - 1 of its partition: wave w renders sub-tile
// st0 + w. Its sub-tiles' lists hold the same Gaussians in the same order
// (the instance kernel writes a Gaussian to all four sub-tiles of a FOVEA or
// BLENDED tile), so the waves walk the list of the first one together, 64 W
// Gaussians a round: lane t of the workgroup loads Gaussian b0 + t and
// computes, per pixel row of the work tile, one 16-column interval of its
// support ellipse, whose two halves are the row's bits in the two sub-tiles
// beside each other. A Gaussian with a sample in any sub-tile is staged (in
// list order, across the waves' ballots); each wave then takes the passes
// over the staged entries with its own sub-tile's masks, 64 at a time.
template <int K, int W>
__device__ __forceinline__ void blend_shared(const ViewParams& vp, const BlendArgs& a, int type, uint32_t j0,
                                             float4* lds) {
    constexpr uint32_t kN = 64u * W;                   // Gaussians per round
    constexpr uint32_t kRows = W == 4 ? 16u : 8u;      // pixel rows of the work tile
    float4* sV1 = lds;
    float4* sV2 = sV1 + kN;
    float4* sV4 = sV2 + kN;
    float4* sC = sV4 + kN;
    uint32_t* sM = reinterpret_cast<uint32_t*>(sC + kN);        // [sub-tile][entry][2]
    uint64_t* sBal = reinterpret_cast<uint64_t*>(sM + 2u * W * kN);  // [round parity][wave]
    const uint32_t tid = threadIdx.x, w = tid >> 6, li = tid & 63u;
    const uint32_t tg0 = a.tile_map_part[j0];
    const uint32_t lt = tg0 / kNumSub, st0 = tg0 % kNumSub;  // st0: 0, or 0 / 2 for W = 2
    const uint32_t ltx = lt % (uint32_t)vp.gw, lty = lt / (uint32_t)vp.gw;
    const uint32_t wx = ltx * kTileL, wy = lty * kTileL + (st0 / kStride) * kTileS;  // work tile origin
    const uint32_t st = st0 + w;  // this wave's sub-tile
    const uint32_t bx = wx + (st % kStride) * kTileS, by = lty * kTileL + (st / kStride) * kTileS;
    const uint32_t px = bx + (li & 7u), py = by + (li >> 3);
    const bool inside = px < (uint32_t)vp.width && py < (uint32_t)vp.height;
    const float pxf = (float)px, pyf = (float)py;
    const uint32_t m_word = li >> 5, m_shift = li & 31u;
    const uint32_t* mine = sM + 2u * kN * w;  // this wave's sub-tile masks

    PixelState<K> ps;
    pixel_init(ps);

    const uint2 range = a.ranges[tg0];
    uint32_t parity = 0u;
    for (uint32_t b0 = range.x; b0 < range.y; b0 += kN) {  // the same rounds in every wave
        const uint32_t idx = b0 + tid;
        float4 V1 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), V2 = V1, V4 = V1, C = V1;
        // m[2 sy + sx]: the mask of sub-tile (sx, sy) of the work tile
        uint32_t m[W][2];
#pragma unroll
        for (int s = 0; s < W; ++s) m[s][0] = m[s][1] = 0u;
        if (idx < range.y) {
            const uint32_t p = a.vals[idx];
            const float4 E = a.ell[2 * (size_t)p];
            const float er[5] = {E.x, E.y, E.z, E.w, a.ell[2 * (size_t)p + 1].x};
            const float4* rec = a.prim + 4 * (size_t)p;
            V1 = rec[0];
            V2 = rec[1];
            V4 = rec[2];
            C = rec[3];
            const FhMaskPrep mp = fh_mask_prep(er, rcp_rn(er[3]), (float)wx, (float)wy, 1.0f);
            if (mp.all) {
#pragma unroll
                for (int s = 0; s < W; ++s) m[s][0] = m[s][1] = ~0u;
            } else {
#pragma unroll
                for (uint32_t r = 0; r < kRows; ++r) {
                    const float dy = (float)r + mp.dy0;
                    const float w2 = fh_mask_row_w2(mp, dy);
                    const float wd = w2 < 0.0f ? -1.0f : __builtin_amdgcn_sqrtf(w2) + kMaskMarginPx;
                    const uint32_t bits = fh_mask_row_bits(mp, dy, wd, 16);
                    const uint32_t sy = r >> 3, rr = r & 7u, sh = (rr & 3u) * 8u;
                    m[2u * sy][rr >> 2] |= (bits & 0xffu) << sh;
                    m[2u * sy + 1u][rr >> 2] |= (bits >> 8) << sh;
                }
            }
        }
        bool keep = false;
#pragma unroll
        for (int s = 0; s < W; ++s) keep |= (m[s][0] | m[s][1]) != 0u;
        const uint64_t kb = __ballot(keep);
        if (li == 0u) sBal[parity * W + w] = kb;
        // every wave's ballot is in; every wave has finished reading the
        // previous round's entries (the ballots alternate between two rows, so
        // a wave a round ahead never overwrites a row another still reads)
        __syncthreads();
        uint32_t total = 0u, pre = 0u;
#pragma unroll
        for (uint32_t v = 0; v < (uint32_t)W; ++v) {
            const uint32_t c = (uint32_t)__popcll(sBal[parity * W + v]);
            pre += v < w ? c : 0u;
            total += c;
        }
        parity ^= 1u;
        if (total == 0u) continue;  // the same in every wave
        const uint32_t pos = pre + (uint32_t)__popcll(kb & lanemask_lt());
        if (keep) {
            sV1[pos] = V1;
            sV2[pos] = V2;
            sV4[pos] = V4;
            sC[pos] = C;
#pragma unroll
            for (int s = 0; s < W; ++s) {
                sM[2u * (kN * s + pos)] = m[s][0];
                sM[2u * (kN * s + pos) + 1u] = m[s][1];
            }
        }
        __syncthreads();
        if (!inside) continue;
        for (uint32_t e0 = 0; e0 < total; e0 += 64u)
            blend_batch<K>(ps, mine, e0, (int)min(64u, total - e0), m_word, m_shift, sV1, sV2, sV4, sC, pxf, pyf);
    }
    if (!inside) return;
    pixel_resolve_store<K>(ps, vp, type, px, py, pxf, pyf);
}

// 16x8 (W = 2) or 16x16 (W = 4) work tiles: one W-wave workgroup per unit of
// the work order (wide_unit): a FOVEA or BLENDED work tile staged once for
// its W sub-tiles, or up to W PERIPHERY tiles (one wave each) or up to 4 W
// QUARTER tiles (four per wave), whose waves share nothing.
template <int K, int W>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(K <= 8 ? 5 : 1))) void blend_wide_kernel(const ViewParams* vpp, BlendArgs a) {
    const ViewParams vp = load_view(vpp);
    __shared__ float4 lds[wide_lds_f4<W>()];
    const WideUnits U = wide_units(a.frame_info, a.n_small, (uint32_t)W);
    const uint32_t g0 = xcd_remap(blockIdx.x, gridDim.x, a.xcd_run);
    if (g0 >= U.end) return;
    const WideUnit wu = wide_unit(U, a.order[g0], (uint32_t)W);
    if (wu.type == kTileFovea || wu.type == kTileBlended) {
        blend_shared<K, W>(vp, a, wu.type, wu.first, lds);
        return;
    }
    const uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    float4* base = lds + w * kWaveLdsF4;
    uint32_t* sM = reinterpret_cast<uint32_t*>(base + 256);
    if (wu.type == kTilePeriphery) {
        if (w < wu.count)
            blend_wave<K, true>(vp, a, kTilePeriphery, wu.first + w, true, lane, base, base + 64, base + 128,
                                base + 192, sM);
    } else if (4u * w < wu.count) {
        const uint32_t j = wu.first + 4u * w + (lane >> 4);
        blend_wave<K, true>(vp, a, kTileQuarter, j, j < wu.first + wu.count, lane, base, base + 64, base + 128,
                            base + 192, sM);
    }
}

// Blend work order. A work unit is one active FOVEA, PERIPHERY or BLENDED
// tile (partition index u) or four QUARTER tiles (unit u >= the non-QUARTER
// count, in partition order). A unit's cost is its longest instance list;
// units are bucketed by cost, four buckets per octave, and listed heaviest
// bucket first, so that the longest workgroups start first and the short
// ones fill the end of the launch instead of a few long ones trailing it.
// Two passes over many workgroups (bucket histogram, then placement); the
// order inside a bucket is whatever the atomics give: it decides which
// workgroup renders a unit, never a pixel value.
constexpr int kOrderBuckets = 64;
constexpr int kOrderThreads = 256;
static_assert(kOrderBuckets <= 64, "ViewBuffers::order_counts holds 2 x 64 words");
static_assert(kOrderBuckets <= kOrderThreads, "one thread per bucket clears and flushes the counts");
constexpr uint32_t kOrderMaxBlocks = 512;

// bucket 0 holds the longest lists: 4 floor(log2(c+1)) plus the quarter of
// the octave (the two bits below the leading one)
__device__ __forceinline__ uint32_t cost_bucket(uint32_t c) {
    const uint32_t v = c + 1u;
    const uint32_t l = 31u - (uint32_t)__clz(v);
    const uint32_t qtr = l > 1u ? (v >> (l - 2u)) & 3u : (l == 1u ? (v & 1u) << 1 : 0u);
    return (uint32_t)kOrderBuckets - 1u - min(4u * l + qtr, (uint32_t)kOrderBuckets - 1u);
}

struct OrderUnits {
    uint32_t na, n_nq, units;
};

__device__ __forceinline__ OrderUnits order_units(const uint32_t* frame_info, uint32_t n_small) {
    OrderUnits o;
    o.na = min(frame_info[FI_N_ACTIVE], n_small);
    const uint32_t nq = min(frame_info[FI_PART + 7] - frame_info[FI_PART + 6], o.na);
    o.n_nq = o.na - nq;
    o.units = o.n_nq + (nq + 3u) / 4u;
    return o;
}

// Units of the work order: the single-wave kernel's (W = 1) or wide_units.
template <int W>
__device__ __forceinline__ uint32_t order_unit_count(const uint32_t* frame_info, uint32_t n_small) {
    if constexpr (W == 1) return order_units(frame_info, n_small).units;
    else return wide_units(frame_info, n_small, (uint32_t)W).end;
}

// Pass 1: each unit's bucket, and the bucket histogram (LDS per workgroup,
// then one global add per bucket). W = 1: a non-QUARTER unit u is
// partitioned tile u; QUARTER unit u covers partitioned tiles
// n_nq + 4 (u - n_nq) + k, k < 4. W = 2, 4: wide_unit; a FOVEA or BLENDED
// work tile's cost is its list's length (its sub-tiles share it).
template <int W>
__global__ __launch_bounds__(kOrderThreads) void blend_order_count_kernel(const uint32_t* frame_info,
                                                                          const uint32_t* tile_map_part,
                                                                          const uint2* ranges, uint8_t* bucket,
                                                                          uint32_t* counts, uint32_t n_small) {
    __shared__ uint32_t s_cnt[kOrderBuckets];
    const uint32_t units = order_unit_count<W>(frame_info, n_small);
    const OrderUnits o = order_units(frame_info, n_small);
    const WideUnits U = wide_units(frame_info, n_small, (uint32_t)W);
    if (threadIdx.x < (uint32_t)kOrderBuckets) s_cnt[threadIdx.x] = 0u;
    __syncthreads();
    for (uint32_t u = blockIdx.x * kOrderThreads + threadIdx.x; u < units; u += gridDim.x * kOrderThreads) {
        uint32_t c = 0u;
        if constexpr (W == 1) {
            const bool qu = u >= o.n_nq;
            const uint32_t j0 = qu ? o.n_nq + (u - o.n_nq) * 4u : u;
            const uint32_t cnt = qu ? min(4u, o.na - j0) : 1u;
            uint32_t tg[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) tg[k] = (uint32_t)k < cnt ? tile_map_part[j0 + k] : 0u;
#pragma unroll
            for (int k = 0; k < 4; ++k) {  // unconditional loads (tile 0 stands in for none)
                const uint2 r = ranges[tg[k]];
                c = (uint32_t)k < cnt ? max(c, r.y - r.x) : c;
            }
        } else {
            const WideUnit wu = wide_unit(U, u, (uint32_t)W);
            const uint32_t cnt = (wu.type == kTileFovea || wu.type == kTileBlended) ? 1u : wu.count;
            for (uint32_t k = 0; k < cnt; ++k) {
                const uint2 r = ranges[tile_map_part[wu.first + k]];
                c = max(c, r.y - r.x);
            }
        }
        const uint32_t b = cost_bucket(c);
        bucket[u] = (uint8_t)b;
        atomicAdd(&s_cnt[b], 1u);
    }
    __syncthreads();
    if (threadIdx.x < (uint32_t)kOrderBuckets && s_cnt[threadIdx.x] != 0u)
        atomicAdd(&counts[threadIdx.x], s_cnt[threadIdx.x]);
}

// Pass 2: bucket starts from the histogram (every workgroup scans the same
// 48 counts), a range per (workgroup round, bucket) from the global cursors,
// ranks inside it from LDS atomics.
template <int W>
__global__ __launch_bounds__(kOrderThreads) void blend_order_place_kernel(const uint32_t* frame_info,
                                                                          const uint8_t* bucket, uint32_t* counts,
                                                                          uint32_t* order, uint32_t n_small) {
    __shared__ uint32_t s_start[kOrderBuckets], s_cnt[kOrderBuckets], s_base[kOrderBuckets];
    const uint32_t units = order_unit_count<W>(frame_info, n_small);
    if (threadIdx.x == 0) {
        uint32_t run = 0u;
        for (int b = 0; b < kOrderBuckets; ++b) {
            s_start[b] = run;
            run += counts[b];
        }
    }
    for (uint32_t u0 = blockIdx.x * kOrderThreads; u0 < units; u0 += gridDim.x * kOrderThreads) {
        if (threadIdx.x < (uint32_t)kOrderBuckets) s_cnt[threadIdx.x] = 0u;
        __syncthreads();
        const uint32_t u = u0 + threadIdx.x;
        const bool on = u < units;
        const uint32_t b = on ? bucket[u] : 0u;
        const uint32_t r = on ? atomicAdd(&s_cnt[b], 1u) : 0u;
        __syncthreads();
        if (threadIdx.x < (uint32_t)kOrderBuckets && s_cnt[threadIdx.x] != 0u)
            s_base[threadIdx.x] = atomicAdd(&counts[64 + threadIdx.x], s_cnt[threadIdx.x]);
        __syncthreads();
        if (on) order[s_start[b] + s_base[b] + r] = u;
        __syncthreads();
    }
}

template <int W>
void launch_order_w(const ViewBuffers& vb, uint32_t n_small, uint32_t blocks, hipStream_t s) {
    blend_order_count_kernel<W><<<blocks, kOrderThreads, 0, s>>>(vb.frame_info, vb.tile_map_part, vb.ranges,
                                                                  vb.order_bucket, vb.order_counts, n_small);
    blend_order_place_kernel<W><<<blocks, kOrderThreads, 0, s>>>(vb.frame_info, vb.order_bucket,
                                                                  vb.order_counts, vb.blend_order, n_small);
}

hipError_t launch_blend_order(const ViewBuffers& vb, int n_small, int work_tile, hipStream_t s) {
    const uint32_t blocks = std::min<uint32_t>(kOrderMaxBlocks, ((uint32_t)n_small + kOrderThreads - 1u) / kOrderThreads);
    if (work_tile == 4) launch_order_w<4>(vb, (uint32_t)n_small, blocks, s);
    else if (work_tile == 2) launch_order_w<2>(vb, (uint32_t)n_small, blocks, s);
    else launch_order_w<1>(vb, (uint32_t)n_small, blocks, s);
    return hipGetLastError();
}

template <int W>
void launch_blend_w(const ViewParams* dvp, const BlendArgs& a, int K, uint32_t grid, hipStream_t s) {
    // rasterization/src/inference.cu:17-31: largest power of two <= K in 1..32
    constexpr uint32_t T = 64u * W;
    if constexpr (W == 1) {
        if (K >= 32) blend_kernel<32><<<grid, T, 0, s>>>(dvp, a);
        else if (K >= 16) blend_kernel<16><<<grid, T, 0, s>>>(dvp, a);
        else if (K >= 8) blend_kernel<8><<<grid, T, 0, s>>>(dvp, a);
        else if (K >= 4) blend_kernel<4><<<grid, T, 0, s>>>(dvp, a);
        else if (K >= 2) blend_kernel<2><<<grid, T, 0, s>>>(dvp, a);
        else blend_kernel<1><<<grid, T, 0, s>>>(dvp, a);
    } else {
        if (K >= 32) blend_wide_kernel<32, W><<<grid, T, 0, s>>>(dvp, a);
        else if (K >= 16) blend_wide_kernel<16, W><<<grid, T, 0, s>>>(dvp, a);
        else if (K >= 8) blend_wide_kernel<8, W><<<grid, T, 0, s>>>(dvp, a);
        else if (K >= 4) blend_wide_kernel<4, W><<<grid, T, 0, s>>>(dvp, a);
        else if (K >= 2) blend_wide_kernel<2, W><<<grid, T, 0, s>>>(dvp, a);
        else blend_wide_kernel<1, W><<<grid, T, 0, s>>>(dvp, a);
    }
}

// `grid` is the number of work units (blend_units), or a bound on it:
// workgroups past the frame's units return at once.
hipError_t launch_blend(const ViewParams* dvp, const ViewBuffers& vb, const uint32_t* sorted_vals,
                        int K, uint32_t grid, int xcd_run, int work_tile, uint32_t n_small, hipStream_t s) {
    if (grid == 0) return hipSuccess;
    BlendArgs a;
    a.frame_info = vb.frame_info;
    a.order = vb.blend_order;
    a.tile_map_part = vb.tile_map_part;
    a.ranges = vb.ranges;
    a.vals = sorted_vals;
    a.prim = vb.prim;
    a.ell = vb.ell;
    a.xcd_run = (uint32_t)(xcd_run > 0 ? xcd_run : 0);
    a.n_small = n_small;
    if (work_tile == 4) launch_blend_w<4>(dvp, a, K, grid, s);
    else if (work_tile == 2) launch_blend_w<2>(dvp, a, K, grid, s);
    else launch_blend_w<1>(dvp, a, K, grid, s);
    return hipGetLastError();
}

// Work units of a frame with the partition ranges `part` (FI_PART words) and
// n_active active tiles: the grid of an exact blend launch.
uint32_t blend_units(const uint32_t* part, uint32_t n_active, int work_tile) {
    if (work_tile <= 1) {
        const uint32_t nq = part[7] - part[6];
        return n_active - nq + (nq + 3u) / 4u;
    }
    const uint32_t W = (uint32_t)work_tile;
    return (part[1] - part[0]) / W + (part[3] - part[2] + W - 1u) / W + (part[5] - part[4]) / W +
           (part[7] - part[6] + 4u * W - 1u) / (4u * W);
}

}  // namespace fhtgs