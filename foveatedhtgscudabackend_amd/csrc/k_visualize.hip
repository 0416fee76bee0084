// Debug overlays drawn into a rendered image (visualization/include/kernels/
// visualization.cuh, visualization/src/visualization.cu): a gaze marker, and
// the borders of the active tiles coloured by tile type.
#include "fhtgs_device.h"

namespace fhtgs {

__device__ __forceinline__ void put_rgb(float* img, int x, int y, int W, int H, int chw, float r, float g,
                                        float b) {
    const int64_t idx = (int64_t)W * y + x;
    if (chw) {
        const int64_t np = (int64_t)W * H;
        img[idx] = r;
        img[np + idx] = g;
        img[2 * np + idx] = b;
    } else {
        img[3 * idx] = r;
        img[3 * idx + 1] = g;
        img[3 * idx + 2] = b;
    }
}

// visualize_gaze_cu (visualization.cuh:11-50): 8x8 threads per workgroup over
// the marker's square; the offset of a thread is its index minus size/2,
// computed in float and truncated to int, and offsets above size/2 are
// skipped (so an odd size draws offset 0 twice, harmlessly).
__global__ __launch_bounds__(64) void gaze_marker_kernel(float* img, int W, int H, int gx, int gy, float r,
                                                         float g, float b, float size, int type, int chw) {
    const int tx = (int)(threadIdx.x & 7u), ty = (int)(threadIdx.x >> 3);
    const float half = size / 2.0f;
    const int xo = (int)((float)(blockIdx.x * 8u + (unsigned)tx) - half);
    const int yo = (int)((float)(blockIdx.y * 8u + (unsigned)ty) - half);
    if ((float)xo > half || (float)yo > half) return;
    if (type == 0) {  // circle
        if ((float)(xo * xo + yo * yo) > half * half) return;
    } else if (type == 2) {  // cross
        if (abs(xo) > 1 && abs(yo) > 1) return;
    }
    const int x = gx + xo, y = gy + yo;
    if (x < 0 || x >= W || y < 0 || y >= H) return;
    put_rgb(img, x, y, W, H, chw, r, g, b);
}

hipError_t launch_gaze_marker(float* img, int W, int H, float gaze_x, float gaze_y, const float rgb[3], int size,
                              int type, int chw, hipStream_t s) {
    const int n = (size + 7) / 8;
    if (size <= 0 || n <= 0) return hipSuccess;
    // the marker centre: the gaze clamped to the image, truncated
    // (visualization.cu:36-40, visualization.cuh:33-34)
    const float cx = fmaxf(0.0f, fminf(gaze_x, (float)(W - 1)));
    const float cy = fmaxf(0.0f, fminf(gaze_y, (float)(H - 1)));
    gaze_marker_kernel<<<dim3((unsigned)n, (unsigned)n), 64, 0, s>>>(img, W, H, (int)cx, (int)cy, rgb[0], rgb[1],
                                                                      rgb[2], (float)size, type, chw);
    return hipGetLastError();
}

// visualize_tile_boundaries_cu (visualization.cuh:52-110) for all three tile
// types in one launch: workgroup g draws the border of partitioned tile g --
// 8x8 for FOVEA (orange) and BLENDED (light blue) sub-tiles, 16x16 for
// PERIPHERY (green) tiles -- one thread per border pixel. Workgroups past the
// partition ranges (the grid may be a bound) return.
__global__ __launch_bounds__(64) void tile_border_kernel(const ViewParams* vpp, const uint32_t* frame_info,
                                                         const uint32_t* tile_map_part, float* img, int chw) {
    const ViewParams vp = load_view(vpp);
    const uint32_t g = blockIdx.x;
    int type;
    if (g >= frame_info[FI_PART + 0] && g < frame_info[FI_PART + 1]) type = kTileFovea;
    else if (g >= frame_info[FI_PART + 2] && g < frame_info[FI_PART + 3]) type = kTilePeriphery;
    else if (g >= frame_info[FI_PART + 4] && g < frame_info[FI_PART + 5]) type = kTileBlended;
    else return;
    const uint32_t tg = tile_map_part[g];
    const uint32_t lt = tg / kNumSub, st = tg % kNumSub;
    const int w = type == kTilePeriphery ? kTileL : kTileS;
    const int h = w;
    const int i = (int)threadIdx.x;
    if (i >= 2 * w + 2 * h - 4) return;
    int px, py;
    if (i < w) { px = i; py = 0; }
    else if (i < 2 * w) { px = i - w; py = h - 1; }
    else if (i < 2 * w + h - 2) { px = 0; py = i - 2 * w + 1; }
    else { px = w - 1; py = i - (2 * w + h - 2) + 1; }
    const int x = (int)(lt % (uint32_t)vp.gw) * kTileL + (int)(st % kStride) * kTileS + px;
    const int y = (int)(lt / (uint32_t)vp.gw) * kTileL + (int)(st / kStride) * kTileS + py;
    if (x >= vp.width || y >= vp.height) return;
    // Okabe-Ito colours of visualization.cu:183-226
    float r, gr, b;
    if (type == kTileFovea) { r = 0.902f; gr = 0.624f; b = 0.0f; }
    else if (type == kTileBlended) { r = 0.337f; gr = 0.706f; b = 0.914f; }
    else { r = 0.0f; gr = 0.620f; b = 0.451f; }
    put_rgb(img, x, y, vp.width, vp.height, chw, r, gr, b);
}

hipError_t launch_tile_borders(const ViewParams& vp, const ViewParams* dvp, const ViewBuffers& vb, float* img,
                               int chw, hipStream_t s) {
    if (vp.n_small <= 0) return hipSuccess;
    tile_border_kernel<<<(unsigned)vp.n_small, 64, 0, s>>>(dvp, vb.frame_info, vb.tile_map_part, img, chw);
    return hipGetLastError();
}

}  // namespace fhtgs
