// Foveal upsampling in one launch for all PERIPHERY and BLENDED tiles
// (rasterization/include/kernels/interpolation.cuh):
//   * PERIPHERY large tile (workgroup of 256 = 16x16 pixels): bilinear fill of
//     the pixels between the stride-2 samples (interpolation.cuh:41-146), or
//     with blur_periphery the fused [1 6 1]^2 blur (interpolation.cuh:270-413);
//   * BLENDED small tiles (four 8x8 tiles per workgroup): the same upsample
//     cross-faded with the full-resolution render by the ring factor
//     (interpolation.cuh:149-267, 416-584).
// Samples are always even/even pixels, which this launch never writes, so the
// in-place variant has no read/write overlap between workgroups.
// Extension: QUARTER large tiles (one workgroup each) fill their 16x16 pixels
// bilinearly from the 5x5 stride-4 samples (fractions 0, 1/4, 1/2, 3/4) of
// the tile and its right and lower neighbours. A QUARTER tile's even/even
// pixels are not all samples, so the stride-2 grids treat QUARTER tiles like
// invisible ones (clamped into their own tile); stride-4 positions of every
// visible tile are written by the blend only, so the quarter grid reads any
// visible neighbour.
#include "fhtgs_device.h"

namespace fhtgs {

__device__ __forceinline__ v3 read_rgb(const float* img, int x, int y, int W, int H, int chw) {
    const int sx = max(0, min(W - 1, x));
    const int sy = max(0, min(H - 1, y));
    const int64_t idx = (int64_t)W * sy + sx;
    if (chw) {
        const int64_t n = (int64_t)W * H;
        return v3{img[idx], img[n + idx], img[2 * n + idx]};
    }
    return v3{img[3 * idx], img[3 * idx + 1], img[3 * idx + 2]};
}

__device__ __forceinline__ void write_rgb(float* img, int x, int y, int W, int H, int chw, v3 c) {
    const int64_t idx = (int64_t)W * y + x;
    if (chw) {
        const int64_t n = (int64_t)W * H;
        img[idx] = c.x; img[n + idx] = c.y; img[2 * n + idx] = c.z;
    } else {
        img[3 * idx] = c.x; img[3 * idx + 1] = c.y; img[3 * idx + 2] = c.z;
    }
}

// The tile-type rule of the tiles kernel: a visible large tile is QUARTER
// when its centre lies at least quarter_r tiles from the gaze tile.
__device__ __forceinline__ bool quarter_tile(const ViewParams& vp, int tile) {
    return vp.quarter_r > 0 && !in_fovea(vp.quarter_r, (float)(tile % vp.gw) + 0.5f,
                                         (float)(tile / vp.gw) + 0.5f, (float)vp.gtx, (float)vp.gty);
}

// One sample of the stride-S grid starting at (bx, by); samples in invisible
// large tiles -- and, for S = 2, in QUARTER tiles -- are clamped into the
// tile (lx, ly) (interpolation.cuh:70-101).
template <int S>
__device__ __forceinline__ v3 grid_sample(const ViewParams& vp, const float* img, int lx, int ly,
                                          int bx, int by, int sx, int sy) {
    const int mx = (vp.width - bx - 1) / S, my = (vp.height - by - 1) / S;
    const int px = bx + min(sx, mx) * S;
    const int py = by + min(sy, my) * S;
    const int tile = (py / kTileL) * vp.gw + px / kTileL;
    if (!mask_bit(vp.vis_mask, (uint32_t)tile) || (S == kStride && quarter_tile(vp, tile))) {
        const int cx = max(lx * kTileL, min(lx * kTileL + kTileL - S, px));
        const int cy = max(ly * kTileL, min(ly * kTileL + kTileL - S, py));
        return read_rgb(img, cx, cy, vp.width, vp.height, vp.to_chw);
    }
    return read_rgb(img, px, py, vp.width, vp.height, vp.to_chw);
}

__device__ __forceinline__ void acc_fma(v3& a, v3 s, float w) {
    a.x = fmaf(s.x, w, a.x); a.y = fmaf(s.y, w, a.y); a.z = fmaf(s.z, w, a.z);
}
__device__ __forceinline__ void acc_add(v3& a, v3 s) {
    a.x = a.x + s.x; a.y = a.y + s.y; a.z = a.z + s.z;
}

// interpolation.cuh:112-132 (frac in {0, 1/2}); for the stride-4 QUARTER
// grid the same weights with frac in {0, 1/4, 1/2, 3/4}
template <int S>
__device__ __forceinline__ v3 bilinear(const v3* s, int nx, int bsx, int bsy, int xm, int ym) {
    const float xf = (float)xm / (float)S, yf = (float)ym / (float)S;
    v3 acc{0.0f, 0.0f, 0.0f};
    if (xm == 0) {
        acc_fma(acc, s[bsy * nx + bsx], 1.0f - yf);
        acc_fma(acc, s[(bsy + 1) * nx + bsx], yf);
    } else if (ym == 0) {
        acc_fma(acc, s[bsy * nx + bsx], 1.0f - xf);
        acc_fma(acc, s[bsy * nx + bsx + 1], xf);
    } else {
        const v3 a = s[bsy * nx + bsx], b = s[bsy * nx + bsx + 1];
        const v3 c = s[(bsy + 1) * nx + bsx], d = s[(bsy + 1) * nx + bsx + 1];
        acc_fma(acc, v3{a.x * (1.0f - xf), a.y * (1.0f - xf), a.z * (1.0f - xf)}, 1.0f - yf);
        acc_fma(acc, v3{b.x * xf, b.y * xf, b.z * xf}, 1.0f - yf);
        acc_fma(acc, v3{c.x * (1.0f - xf), c.y * (1.0f - xf), c.z * (1.0f - xf)}, yf);
        acc_fma(acc, v3{d.x * xf, d.y * xf, d.z * xf}, yf);
    }
    return acc;
}

// interpolation.cuh:343-399; returns the 1/64, 1/16 or 1/4 factor
__device__ __forceinline__ float blur_sum(const v3* s, int nx, int bsx, int bsy, int tx, int ty, v3& acc) {
    acc = v3{0.0f, 0.0f, 0.0f};
    auto at = [&](int y, int x) { return s[y * nx + x]; };
    if ((tx & 1) == 0 && (ty & 1) == 0) {
        acc_add(acc, at(bsy - 1, bsx - 1)); acc_fma(acc, at(bsy - 1, bsx), 6.0f); acc_add(acc, at(bsy - 1, bsx + 1));
        acc_fma(acc, at(bsy, bsx - 1), 6.0f); acc_fma(acc, at(bsy, bsx), 36.0f); acc_fma(acc, at(bsy, bsx + 1), 6.0f);
        acc_add(acc, at(bsy + 1, bsx - 1)); acc_fma(acc, at(bsy + 1, bsx), 6.0f); acc_add(acc, at(bsy + 1, bsx + 1));
        return 1.0f / 64.0f;
    } else if ((tx & 1) == 0) {
        acc_add(acc, at(bsy, bsx - 1)); acc_fma(acc, at(bsy, bsx), 6.0f); acc_add(acc, at(bsy, bsx + 1));
        acc_add(acc, at(bsy + 1, bsx - 1)); acc_fma(acc, at(bsy + 1, bsx), 6.0f); acc_add(acc, at(bsy + 1, bsx + 1));
        return 1.0f / 16.0f;
    } else if ((ty & 1) == 0) {
        acc_add(acc, at(bsy - 1, bsx)); acc_add(acc, at(bsy - 1, bsx + 1));
        acc_fma(acc, at(bsy, bsx), 6.0f); acc_fma(acc, at(bsy, bsx + 1), 6.0f);
        acc_add(acc, at(bsy + 1, bsx)); acc_add(acc, at(bsy + 1, bsx + 1));
        return 1.0f / 16.0f;
    }
    acc_add(acc, at(bsy, bsx)); acc_add(acc, at(bsy, bsx + 1));
    acc_add(acc, at(bsy + 1, bsx)); acc_add(acc, at(bsy + 1, bsx + 1));
    return 1.0f / 4.0f;
}

// interpolation.cuh:244-250
__device__ __forceinline__ float ring_factor(const ViewParams& vp, int px, int py) {
    const float dx = vp.gcx - (float)px, dy = vp.gcy - (float)py;
    const float len = sqrtf(fmaf(dy, dy, dx * dx));
    return fh_clamp(((len - (float)vp.blend_radius_px) - (float)kTileL) / (float)(vp.blend_width_px - kTileL),
                    0.0f, 1.0f);
}

__device__ __forceinline__ v3 sat3(v3 c) { return v3{fh_saturate(c.x), fh_saturate(c.y), fh_saturate(c.z)}; }

__global__ __launch_bounds__(256) void upsample_kernel(const ViewParams* vpp, const uint32_t* frame_info,
                                                       const uint32_t* tile_map_part) {
    const ViewParams vp = load_view(vpp);
    const uint32_t n_periphery = frame_info[FI_PART + 3] - frame_info[FI_PART + 2];
    __shared__ v3 s_smp[4 * 36 > 100 ? 4 * 36 : 100];
    const int tid = threadIdx.x;
    const int W = vp.width, H = vp.height;
    const float* src = vp.blur ? vp.image_blend : vp.image_final;
    if (blockIdx.x < n_periphery) {
        const uint32_t lt = tile_map_part[frame_info[FI_PART + 2] + blockIdx.x] / kNumSub;
        const int lx = (int)(lt % (uint32_t)vp.gw), ly = (int)(lt / (uint32_t)vp.gw);
        const int ns = vp.blur ? kTileS + 2 : kTileS + 1;
        const int off = vp.blur ? kStride : 0;
        const int bx = lx * kTileL - off, by = ly * kTileL - off;
        for (int k = tid; k < ns * ns; k += 256)
            s_smp[k] = grid_sample<kStride>(vp, src, lx, ly, bx, by, k % ns, k / ns);
        __syncthreads();
        const int tx = tid & 15, ty = tid >> 4;
        const int px = lx * kTileL + tx, py = ly * kTileL + ty;
        if (px >= W || py >= H) return;
        if (!vp.blur) {
            const int xm = tx % kStride, ym = ty % kStride;
            if (xm == 0 && ym == 0) return;
            write_rgb(vp.image_final, px, py, W, H, vp.to_chw,
                      sat3(bilinear<kStride>(s_smp, ns, tx / kStride, ty / kStride, xm, ym)));
        } else {
            v3 acc;
            const float f = blur_sum(s_smp, ns, tx / 2 + 1, ty / 2 + 1, tx, ty, acc);
            write_rgb(vp.image_final, px, py, W, H, vp.to_chw, sat3(v3{acc.x * f, acc.y * f, acc.z * f}));
        }
        return;
    }
    const uint32_t nb = frame_info[FI_PART + 5] - frame_info[FI_PART + 4];
    const uint32_t nb_groups = (nb + 3u) / 4u;
    if (blockIdx.x >= n_periphery + nb_groups) {
        // QUARTER: one large tile per workgroup
        const uint32_t j = blockIdx.x - n_periphery - nb_groups;
        if (j >= frame_info[FI_PART + 7] - frame_info[FI_PART + 6]) return;
        const uint32_t lt = tile_map_part[frame_info[FI_PART + 6] + j] / kNumSub;
        const int lx = (int)(lt % (uint32_t)vp.gw), ly = (int)(lt / (uint32_t)vp.gw);
        constexpr int ns = kTileL / kQuarterStride + 1;
        const int bx = lx * kTileL, by = ly * kTileL;
        for (int k = tid; k < ns * ns; k += 256)
            s_smp[k] = grid_sample<kQuarterStride>(vp, vp.image_final, lx, ly, bx, by, k % ns, k / ns);
        __syncthreads();
        const int tx = tid & 15, ty = tid >> 4;
        const int px = bx + tx, py = by + ty;
        if (px >= W || py >= H) return;
        const int xm = tx % kQuarterStride, ym = ty % kQuarterStride;
        if (xm == 0 && ym == 0) return;
        write_rgb(vp.image_final, px, py, W, H, vp.to_chw,
                  sat3(bilinear<kQuarterStride>(s_smp, ns, tx / kQuarterStride, ty / kQuarterStride, xm, ym)));
        return;
    }
    // BLENDED: four small tiles per workgroup
    const int q = tid >> 6, l = tid & 63;
    const uint32_t j = (blockIdx.x - n_periphery) * 4u + (uint32_t)q;
    const bool have = j < nb;
    const int ns = vp.blur ? kTileS / kStride + 2 : kTileS / kStride + 1;
    v3* smp = s_smp + q * 36;
    int lx = 0, ly = 0, ox = 0, oy = 0;
    if (have) {
        const uint32_t tg = tile_map_part[frame_info[FI_PART + 4] + j];
        const uint32_t lt = tg / kNumSub, st = tg % kNumSub;
        lx = (int)(lt % (uint32_t)vp.gw); ly = (int)(lt / (uint32_t)vp.gw);
        ox = lx * kTileL + (int)(st % kStride) * kTileS;
        oy = ly * kTileL + (int)(st / kStride) * kTileS;
        const int off = vp.blur ? kStride : 0;
        for (int k = l; k < ns * ns; k += 64)
            smp[k] = grid_sample<kStride>(vp, src, lx, ly, ox - off, oy - off, k % ns, k / ns);
    }
    __syncthreads();
    if (!have) return;
    const int tx = l & 7, ty = l >> 3;
    const int px = ox + tx, py = oy + ty;
    if (px >= W || py >= H) return;
    if (!vp.blur) {
        const int xm = tx % kStride, ym = ty % kStride;
        if (xm == 0 && ym == 0) return;
        const v3 ip = bilinear<kStride>(smp, ns, tx / kStride, ty / kStride, xm, ym);
        const float bf = ring_factor(vp, px, py);
        const v3 ctr = read_rgb(vp.image_final, px, py, W, H, vp.to_chw);
        if (bf > 0.0f) {
            const float cb = 1.0f - bf;
            write_rgb(vp.image_final, px, py, W, H, vp.to_chw,
                      sat3(v3{fmaf(ip.x, bf, ctr.x * cb), fmaf(ip.y, bf, ctr.y * cb), fmaf(ip.z, bf, ctr.z * cb)}));
        }
    } else {
        v3 acc;
        const float f = blur_sum(smp, ns, tx / 2 + 1, ty / 2 + 1, tx, ty, acc);
        const float bf = ring_factor(vp, px, py);
        const v3 ctr = read_rgb(vp.image_blend, px, py, W, H, vp.to_chw);
        v3 o;
        if (bf > 0.0f) {
            const float cb = 1.0f - bf;
            o = sat3(v3{fmaf(ctr.x, cb, (acc.x * f) * bf), fmaf(ctr.y, cb, (acc.y * f) * bf),
                        fmaf(ctr.z, cb, (acc.z * f) * bf)});
        } else {
            o = ctr;
        }
        write_rgb(vp.image_final, px, py, W, H, vp.to_chw, o);
    }
}

// One workgroup per PERIPHERY large tile, then one per four BLENDED small
// tiles (a blended large tile's four), then one per QUARTER large tile;
// `grid` is that count, or n_large as a bound -- workgroups past the
// partition find no tile and return.
hipError_t launch_upsample(const ViewParams* dvp, const ViewBuffers& vb, uint32_t grid, hipStream_t s) {
    if (grid == 0) return hipSuccess;
    upsample_kernel<<<grid, 256, 0, s>>>(dvp, vb.frame_info, vb.tile_map_part);
    return hipGetLastError();
}

}  // namespace fhtgs
