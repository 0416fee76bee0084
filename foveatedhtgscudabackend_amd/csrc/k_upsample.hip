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

// One stride-2 sample of the grid starting at (bx, by); samples in invisible
// large tiles are clamped into the tile (lx, ly) (interpolation.cuh:70-101).
__device__ __forceinline__ v3 grid_sample(const ViewParams& vp, const float* img, int lx, int ly,
                                          int bx, int by, int sx, int sy) {
    const int mx = (vp.width - bx - 1) / kStride, my = (vp.height - by - 1) / kStride;
    const int px = bx + min(sx, mx) * kStride;
    const int py = by + min(sy, my) * kStride;
    const int tile = (py / kTileL) * vp.gw + px / kTileL;
    if (!mask_bit(vp.vis_mask, (uint32_t)tile)) {
        const int cx = max(lx * kTileL, min(lx * kTileL + kTileL - kStride, px));
        const int cy = max(ly * kTileL, min(ly * kTileL + kTileL - kStride, py));
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

// interpolation.cuh:112-132 (frac in {0, 1/2})
__device__ __forceinline__ v3 bilinear(const v3* s, int nx, int bsx, int bsy, int xm, int ym) {
    const float xf = (float)xm / 2.0f, yf = (float)ym / 2.0f;
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
            s_smp[k] = grid_sample(vp, src, lx, ly, bx, by, k % ns, k / ns);
        __syncthreads();
        const int tx = tid & 15, ty = tid >> 4;
        const int px = lx * kTileL + tx, py = ly * kTileL + ty;
        if (px >= W || py >= H) return;
        if (!vp.blur) {
            const int xm = tx % kStride, ym = ty % kStride;
            if (xm == 0 && ym == 0) return;
            write_rgb(vp.image_final, px, py, W, H, vp.to_chw,
                      sat3(bilinear(s_smp, ns, tx / kStride, ty / kStride, xm, ym)));
        } else {
            v3 acc;
            const float f = blur_sum(s_smp, ns, tx / 2 + 1, ty / 2 + 1, tx, ty, acc);
            write_rgb(vp.image_final, px, py, W, H, vp.to_chw, sat3(v3{acc.x * f, acc.y * f, acc.z * f}));
        }
        return;
    }
    // BLENDED: four small tiles per workgroup
    const uint32_t nb = frame_info[FI_PART + 5] - frame_info[FI_PART + 4];
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
            smp[k] = grid_sample(vp, src, lx, ly, ox - off, oy - off, k % ns, k / ns);
    }
    __syncthreads();
    if (!have) return;
    const int tx = l & 7, ty = l >> 3;
    const int px = ox + tx, py = oy + ty;
    if (px >= W || py >= H) return;
    if (!vp.blur) {
        const int xm = tx % kStride, ym = ty % kStride;
        if (xm == 0 && ym == 0) return;
        const v3 ip = bilinear(smp, ns, tx / kStride, ty / kStride, xm, ym);
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
// tiles (a blended large tile's four); `grid` is that count, or n_large as a
// bound -- workgroups past the partition find no tile and return.
hipError_t launch_upsample(const ViewParams* dvp, const ViewBuffers& vb, uint32_t grid, hipStream_t s) {
    if (grid == 0) return hipSuccess;
    upsample_kernel<<<grid, 256, 0, s>>>(dvp, vb.frame_info, vb.tile_map_part);
    return hipGetLastError();
}

}  // namespace fhtgs
