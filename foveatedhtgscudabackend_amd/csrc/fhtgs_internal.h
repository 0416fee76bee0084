// Internal structures shared by the host pipeline (fhtgs_api.cpp) and the
// gfx950 kernels. Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fhtgs {

// rasterization/include/config.h:13-45
constexpr float kTransmittanceThreshold = 1e-4f;
constexpr float kMaxFragmentAlpha = 1.0f;
constexpr float kMinAlphaRcp = 255.0f;
constexpr float kMinAlpha = 1.0f / 255.0f;
constexpr float kMaxCutoffSq = 11.0825270903f;
constexpr float kMinAlphaCore = 1.0f / 20.0f;
constexpr int kTileL = 16;
constexpr int kTileS = 8;
constexpr int kStride = 2;
constexpr int kNumSub = 4;
constexpr int kTileFovea = 0;
constexpr int kTilePeriphery = 1;
constexpr int kTileBlended = 2;
// Extension: PERIPHERY tiles beyond the quarter radius are rendered at stride
// 4 (4x4 samples per 16x16 tile, four tiles per wave) and upsampled 4x.
constexpr int kTileQuarter = 3;
constexpr int kQuarterStride = 4;
// Largest 16-px tile grid a view may have (4096 x 4096 pixels): the tile
// classification runs in one workgroup and the instance kernel stages the
// visibility bits of every tile in LDS.
constexpr int kMaxLargeTiles = 65536;

// Sort geometry: 256 threads x 24 keys per onesweep tile. With three
// workgroups per CU (137 VGPRs) the 4.8 M keys of one eye of a 2064x2208
// stereo frame are 785 tiles, about one tile per resident workgroup, and the
// look-back chain is a third shorter than with 16 keys.
constexpr int kSortThreads = 256;
constexpr int kSortItems = 24;
constexpr int kSortTile = kSortThreads * kSortItems;
constexpr int kMaxPasses = 4;
constexpr int kMaxBins = 2048;
constexpr uint32_t kFlagAggregate = 1u << 30;
constexpr uint32_t kFlagInclusive = 2u << 30;
constexpr uint32_t kValueMask = (1u << 30) - 1u;
// Upper bound on look-back polls of one predecessor status word; reaching it
// means a predecessor tile never published, and the pass flags an error and
// moves on instead of waiting forever.
constexpr uint32_t kMaxSpin = 1u << 24;
constexpr uint32_t kSortErrorBit = 2u;

constexpr int kPrimBlock = 256;  // preprocess / instance block (shared partition)

// Frame-info words (device, one block per view), read back once per frame.
enum FrameInfo : int {
    FI_PART = 0,          // 8 words: (x,y) per tile type, fovea, periphery, blended, quarter
    FI_N_ACTIVE = 8,
    FI_INSTANCES = 9,     // I = sum of tiles-touched
    FI_VALID = 10,        // non-sentinel instances (atomic)
    FI_ERROR = 11,        // bit 0: tiles-touched under-count, bit 1: sort inconsistency,
                          // bit 2: instance capacity exceeded, bit 3: culling tables that
                          // do not fit the view (impossible tiles-touched)
    FI_N_VISIBLE = 12,    // preprocess survivors (summed by the block scan)
    FI_WORDS = 16
};
// ViewBuffers::block_vis, bit 31: a Gaussian of the block had a tiles-touched
// the culling tables cannot give (FI_ERROR bit 3 once the block scan has run)
constexpr uint32_t kVisRefused = 0x80000000u;

struct ViewParams {
    const float* M;
    const float* VPM;
    const float* VPR_inv;
    const float* cam;
    const uint32_t* vis_mask;
    const uint32_t* vis_sat;
    const uint32_t* fovea_sat;
    const float* bg_data;
    int bg_type, bg_w, bg_h;
    int width, height, gw, gh, n_large, n_small;
    int fovea_r, blend_r, blend_radius_px, blend_width_px;
    int quarter_r;        // quarter-resolution radius in large tiles; 0 = no quarter level
    int gtx, gty;
    float gcx, gcy;
    int active_sh, total_sh;
    float fx, fy, cx, cy, near_plane, far_plane, scale_mod;
    int to_chw, blur, aaa;
    float* image_final;   // fovea tiles and all upsampled output
    float* image_blend;   // periphery / blended blend target (temp image when blurring)
};

// Scene tensors of one preprocess launch, passed by value in its kernel
// arguments (the preprocess is not part of a captured graph).
struct SceneArgs {
    const float* pos;
    const float* scl;
    const float* rot;
    const float* opa;
    const float* sh0;
    const float* shr;
};

// Per-view device buffers for one frame.
struct ViewBuffers {
    uint32_t* frame_info;     // FI_WORDS
    uint32_t* tile_num;       // n_large
    uint32_t* tile_map;       // n_small (compacted, raster order)
    uint8_t* tile_type;       // n_small
    uint32_t* tile_map_part;  // n_small (partitioned by type)
    uint2* ranges;            // n_small
    uint32_t* blend_order;    // n_small: blend work units, longest instance list first
    uint8_t* order_bucket;    // n_small: cost bucket per work unit
    uint32_t* order_counts;   // 2 * 64: bucket histogram, then placement cursors (zeroed per frame)
    uint32_t* n_touched;      // N
    ushort4* bounds;          // N
    float4* prim;             // 4N: one 64-B record per primitive (VPMT1, VPMT2, VPMT4, rgba)
    float4* ell;              // 2N: support-ellipse record (fh_ellipse_record), 32 B per Gaussian:
                              // (cx, cy, m11, m22), (m12, 0, 0, 0) -- one memory sector per gather
    uint32_t* block_sums;     // ceil(N / kPrimBlock), then exclusive offsets
    uint32_t* block_vis;      // ceil(N / kPrimBlock): visible Gaussians, bit 31 a refused table
    // culling tables built on the device when the view brings none
    uint32_t* tab_mask;       // ceil(n_large / 32)
    uint32_t* tab_vis_sat;    // (gw+1)*(gh+1)
    uint32_t* tab_fovea_sat;  // (2R+1)^2
};

struct SortPlan {
    int key_bytes;   // 2 or 4
    int end_bit;
    int passes;
    int bits[kMaxPasses];
    int shift[kMaxPasses];
    int bins_max;    // max over passes of 1 << bits
};

struct InstanceBuffers {
    void* keys[2];
    uint32_t* vals[2];
    uint32_t* hist;      // kMaxPasses * kMaxBins
    uint32_t* status;    // passes * tiles * bins_max
    uint32_t* tickets;   // kMaxPasses tickets, then one error word
    int64_t capacity;    // instances; the instance, sort and range kernels stop
                         // (FI_ERROR bit 2) when a frame needs more
    int64_t tiles_cap;   // onesweep tiles
};

// ---- launchers (one translation unit per kernel family) ----
// Kernels read their view parameters through `dvp`, a device copy of the
// host ViewParams, and the frame's counts (instances, active tiles, tile-type
// partition) from the frame-info words; the host only sizes the grids, either
// exactly (after reading the counts back) or to the buffers' capacity.
hipError_t launch_clear_frame(const ViewParams& vp, const ViewBuffers& vb, hipStream_t s);
// The same for n_views views, as few launches as possible.
hipError_t launch_clear_frames(const ViewParams* const* vps, const ViewBuffers* const* vbs, int n_views,
                               hipStream_t s);
hipError_t launch_tiles(const ViewParams& vp, const ViewParams* dvp, const ViewBuffers& vb, hipStream_t s);
// The views' parameters and the scene go by value into the kernel arguments.
hipError_t launch_preprocess(const ViewParams* const* vps, const ViewBuffers* vbs, int n_views,
                             const SceneArgs& scene, uint32_t n, hipStream_t s);
// The block sums of n_views views, one workgroup each, in one launch.
hipError_t launch_scan_blocks(const ViewBuffers* const* vbs, const uint32_t* n_blocks, int n_views,
                              hipStream_t s);
hipError_t launch_sort_clear(const InstanceBuffers& ib, const SortPlan& plan, int64_t tiles,
                             hipStream_t s);
hipError_t launch_instances(const ViewParams* dvp, const ViewBuffers& vb, const InstanceBuffers& ib,
                            const SortPlan& plan, uint32_t n, int n_large, hipStream_t s);
hipError_t launch_sort_pass(const InstanceBuffers& ib, const SortPlan& plan, int pass, int src,
                            int64_t n_tiles, const uint32_t* frame_info, uint32_t n_explicit,
                            hipStream_t s);
hipError_t launch_sort_histogram(const InstanceBuffers& ib, const SortPlan& plan, int src,
                                 uint32_t n, hipStream_t s);
hipError_t launch_ranges(const ViewBuffers& vb, const InstanceBuffers& ib, const SortPlan& plan,
                         int src, uint32_t n_grid, hipStream_t s);
// work_tile: waves per blend workgroup, 1 (8x8 work tiles), 2 (16x8) or 4
// (16x16); the work order and the blend launch must use the same value.
hipError_t launch_blend(const ViewParams* dvp, const ViewBuffers& vb, const uint32_t* sorted_vals,
                        int K, uint32_t grid, int xcd_run, int work_tile, uint32_t n_small, hipStream_t s);
hipError_t launch_blend_order(const ViewBuffers& vb, int n_small, int work_tile, hipStream_t s);
// Blend work units of a frame (the grid of an exact launch) from its
// partition ranges (FI_PART words) and active tile count.
uint32_t blend_units(const uint32_t* part, uint32_t n_active, int work_tile);
hipError_t launch_upsample(const ViewParams* dvp, const ViewBuffers& vb, uint32_t grid, hipStream_t s);
hipError_t launch_clear_image(const ViewParams& vp, const ViewParams* dvp, const ViewBuffers& vb,
                              hipStream_t s);
hipError_t launch_widen_keys(const void* keys, int key_bytes, uint32_t* out, uint32_t n,
                             hipStream_t s);
// Culling tables on the device. pack: build the mask `words` from `visible`
// (gh*gw bytes, null = all visible); otherwise `words` is the input mask.
// vis_sat / fovea_sat may be null.
hipError_t launch_build_tables(const uint8_t* visible, bool pack, int gw, int gh, int R, uint32_t* words,
                               uint32_t* vis_sat, uint32_t* fovea_sat, hipStream_t s);
hipError_t launch_gaze_marker(float* img, int W, int H, float gaze_x, float gaze_y, const float rgb[3], int size,
                              int type, int chw, hipStream_t s);
hipError_t launch_tile_borders(const ViewParams& vp, const ViewParams* dvp, const ViewBuffers& vb, float* img,
                               int chw, hipStream_t s);

}  // namespace fhtgs
