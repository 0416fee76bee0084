/*
 * fhtgs.h - C ABI of the MI355X (gfx950) foveated hybrid-transparency
 * Gaussian-splatting rasterizer (libfhtgs.so).
 *
 * Plain C: pointers, sizes and status codes only. Device pointers are HIP
 * device pointers on the handle's device; `stream` is a hipStream_t (NULL =
 * the default stream). Every entry point returns FHTGS_OK (0) or a negative
 * FHTGS_E_* code; fhtgs_last_error() gives the message of the last failure.
 *
 * Reference interfaces replaced (paths relative to FoveatedHTGSCudaBackend/):
 *   fhtgs_render (one view)   <- htgs_foveated::rasterization::inference_wrapper,
 *                                rasterization/include/inference_api.h:8-38 and
 *                                rasterization/src/inference_api.cu:13-122
 *                                (pybind name _C.render, torch_bindings/bindings.cpp:8);
 *                                with clear_output = 0 it replaces fast_inference_wrapper,
 *                                rasterization/include/fast_inference_api.h:8-38
 *                                (_C_benchmarking.benchmark, bindings_benchmarking.cpp:7)
 *   fhtgs_render (two views,  <- inference_stereo_wrapper,
 *     shared preprocess)         rasterization/include/inference_stereo_api.h:7-61,
 *                                rasterization/src/inference_stereo_api.cu:13-192
 *                                (_C_stereo.render, bindings_stereo.cpp:8)
 *   fhtgs_render_batch        <- no reference counterpart: many views of several
 *                                resident scenes per call (the reference renders one
 *                                view or one stereo pair per call)
 *   fhtgs_autotune, fhtgs_tuned_digit_bits, fhtgs_tuned_work_tile, fhtgs_get_option:
 *                                no reference counterpart (the reference fixes its tile
 *                                sizes and digit width at compile time,
 *                                rasterization/include/config.h:21-27)
 *   fhtgs_visualize_gaze      <- htgs_foveated::visualization::visualize_gaze_wrapper,
 *                                visualization/include/visualization_api.h:8-14,
 *                                visualization/src/visualization_api.cu:13-35
 *                                (_C_visualization.visualize_gaze, bindings_visualization.cpp:8)
 *   fhtgs_visualize_tile_boundaries <- visualize_tile_boundaries_wrapper,
 *                                visualization_api.h:16-20, visualization_api.cu:37-63
 *                                (_C_visualization.visualize_tile_boundaries, bindings_visualization.cpp:9)
 *   fhtgs_workspace_query     <- required<T>(P, ...) over the from_blob layouts
 *                                (rasterization/include/utils/buffer_utils.h:17-22,
 *                                36-143): scratch bytes from element counts
 *   fhtgs_workspace_bytes     <- the per-call scratch sizing of required<>/from_blob,
 *                                rasterization/include/utils/buffer_utils.h:10-143
 *   fhtgs_sort_pairs          <- cub::DeviceRadixSort::SortPairs over bits [0,end_bit),
 *                                rasterization/src/inference.cu:267-274
 *   fhtgs_sort_pairs_noalloc  <- no reference counterpart: the same sort for callers
 *                                that cannot allocate (fixed per-device scratch)
 *   fhtgs_build_tables        <- the caller-side culling tables the reference takes as
 *                                arguments (torch_bindings/rasterization.py:28-30;
 *                                layouts read by kernel_utils.cuh:168-210)
 *   fhtgs_fovea_radius_tiles  <- config::foveation_radius_tiles / blend_radius_tiles,
 *                                rasterization/include/config.h:30-35 (made runtime)
 *   fhtgs_debug_read / fhtgs_debug_counts: stage outputs of the last render
 *                                (no reference counterpart; parity tests only)
 */
#ifndef FHTGS_H
#define FHTGS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FHTGS_OK 0
#define FHTGS_E_INVALID (-1)      /* bad argument (shape, pointer, value) */
#define FHTGS_E_HIP (-2)          /* HIP runtime error */
#define FHTGS_E_OUT_OF_MEMORY (-3)
#define FHTGS_E_TABLES (-4)       /* culling tables under-count tiles-touched */
#define FHTGS_E_UNSUPPORTED (-5)

/* Background models: rasterization/include/utils/rasterization_utils.h:5-10 */
#define FHTGS_BACKGROUND_NONE 0
#define FHTGS_BACKGROUND_SH 1          /* 16 x float3 sky SH coefficients */
#define FHTGS_BACKGROUND_TEXTURE 2     /* (3,h,w) equirectangular, activated */
#define FHTGS_BACKGROUND_TEXTURE_RAW 3 /* (3,h,w), sigmoid applied after lookup */

/* Tile types: rasterization/include/utils/enums.h:8-13 */
#define FHTGS_TILE_FOVEA 0
#define FHTGS_TILE_PERIPHERY 1
#define FHTGS_TILE_BLENDED 2

/* Reference foveation constants (rasterization/include/config.h:30-32). */
#define FHTGS_DEFAULT_FOVEATION_RADIUS_PX 375
#define FHTGS_DEFAULT_BLEND_WIDTH_PX 125
#define FHTGS_FOVEATION_OFF (-1)

typedef struct fhtgs_context* fhtgs_handle;

/* Scene tensors, all float32 device memory, C-contiguous
 * (torch_bindings/rasterization.py:73-98 argument order). */
typedef struct {
    const float* positions;  /* (N,3) */
    const float* scales;     /* (N,3) linear standard deviations */
    const float* rotations;  /* (N,4) quaternion (w,x,y,z), used as given */
    const float* opacities;  /* (N) or (N,1), activated */
    const float* sh_0;       /* (N,1,3) */
    const float* sh_rest;    /* (N,S,3) */
    int64_t n_primitives;    /* N */
    int32_t total_sh_bases;  /* S = sh_rest.size(1) */
    int32_t reserved;
} fhtgs_scene;

/* One view: the 21 RasterizerSettings fields (torch_bindings/rasterization.py:22-68)
 * plus the flags and the output buffers the torch wrapper allocates. */
typedef struct {
    const float* M;            /* device, >= 3 rows of 4, row major (rows 0..2 used) */
    const float* VPM;          /* device, 4x4 row major, world -> screen */
    const float* VPR_inv;      /* device, 4x4 row major, screen -> world directions */
    const float* cam_position; /* device, 3 */
    float gaze_x, gaze_y;      /* host values, pixels */
    const uint32_t* visibility_mask;            /* device, ceil(n_large/32) words; NULL with both
                                                   tables NULL: all tiles visible */
    const uint32_t* visibility_mask_area_table; /* device, (gw+1)*(gh+1); NULL (with the next
                                                   also NULL): built in the frame */
    const uint32_t* fovea_mask_area_table;      /* device, (2R+1)^2; NULL: built in the frame */
    const float* background_model_data;         /* device; SH: 48 floats, texture: (3,h,w) */
    int32_t background_model_type;
    int32_t background_width;  /* texture w */
    int32_t background_height; /* texture h */
    int32_t K;
    int32_t active_sh_bases; /* 1 / <= 4 / <= 9 / more: SH degree 0..3, which needs
                                total_sh_bases >= 0 / 3 / 8 / 15 (else FHTGS_E_INVALID) */
    int32_t width;
    int32_t height;
    float focal_x, focal_y, center_x, center_y;
    float near_plane, far_plane, scale_modifier;
    int32_t to_chw;
    int32_t blur_periphery;
    int32_t anti_aliasing;
    int32_t foveation_radius_px; /* FHTGS_DEFAULT_FOVEATION_RADIUS_PX; FHTGS_FOVEATION_OFF */
    int32_t blend_width_px;      /* FHTGS_DEFAULT_BLEND_WIDTH_PX */
    int32_t clear_output;        /* 1: pixels no tile writes read 0 (render); 0: left as is (benchmark) */
    float* image;                /* device float32, 3*H*W, layout by to_chw */
    float* image_temp;           /* device float32, 3*H*W; required iff blur_periphery */
    /* Extension (no reference counterpart): tiles whose centre lies at least
     * this many pixels (rounded up to 16-px tiles) from the gaze tile are
     * rendered at quarter resolution and upsampled 4x; <= 0 = off (the
     * reference's two levels). Must be >= foveation_radius_px; not with
     * blur_periphery. */
    int32_t quarter_radius_px;
    int32_t reserved;
} fhtgs_view;

int fhtgs_create(int device, fhtgs_handle* out);
int fhtgs_destroy(fhtgs_handle h);
const char* fhtgs_last_error(fhtgs_handle h);
const char* fhtgs_version(void);

/* Checks a view the way fhtgs_render does before it enqueues anything, on
 * the host only (no device call, no handle): FHTGS_E_INVALID for bad
 * arguments (including blend_width_px <= 16 with foveation on: the ring's
 * cross-fade divides by blend_width_px - 16, interpolation.cuh:244-250),
 * FHTGS_E_UNSUPPORTED for images of more than 65536 16-px tiles (above
 * 4096 x 4096 pixels; the reference has no such limit, inference.cu:89-105).
 * msg (may be NULL) receives the reason. No reference counterpart: the
 * reference validates nothing. */
int fhtgs_validate_view(const fhtgs_view* view, char* msg, int32_t msg_len);

/* Renders n_views views of one scene, enqueued on `stream`. With
 * share_preprocess != 0 the Gaussians are read once for both views (stereo,
 * n_views <= 2); the views must then agree on active_sh_bases, near/far,
 * scale_modifier and anti_aliasing (the reference stereo path uses the left
 * eye's values). The call returns after the frame is finished (as the
 * reference call does, rasterization/src/inference.cu:468-470). How it waits
 * depends on FHTGS_OPT_FRAME_MODE; the images are the same in every mode. */
int fhtgs_render(fhtgs_handle h, const fhtgs_scene* scene, const fhtgs_view* views,
                 int32_t n_views, int32_t share_preprocess, void* stream);

/* Renders n_views views of n_scenes scenes in one call: view v shows
 * scenes[view_scene[v]] and uses the handle's slot view_slot[v] (NULL: slot
 * v) -- a slot keeps its buffers and instance capacity between calls, so a
 * viewer that always uses the same slots keeps its sort buffers resident.
 * Slots of one call must differ (0 <= slot < 4096). Up to four views of one
 * scene that agree on the preprocess parameters share one preprocess pass,
 * which reads the Gaussians once for all of them (each view is grouped with
 * the next ungrouped such views of the call, adjacent or not: stereo pairs,
 * or views of one scene anywhere in the batch). Each view's image equals
 * what fhtgs_render gives for it alone.
 * fhtgs_debug_read/counts address slots. */
int fhtgs_render_batch(fhtgs_handle h, const fhtgs_scene* scenes, int32_t n_scenes,
                       const fhtgs_view* views, const int32_t* view_scene,
                       const int32_t* view_slot, int32_t n_views, void* stream);

/* Autotuner for a batch (no reference counterpart): renders it once exactly,
 * then times device-count frames (best of `frames`) to pick, per view size,
 * the sort's digit width (8 or 11 bits) and the blend's work tile (8x8, 16x8
 * or 16x16 pixels rendered by one workgroup from one staged list,
 * FHTGS_OPT_BLEND_WORK_TILE), and, for the batch, the blend's XCD run (0, 16
 * or 64) and the number of concurrent view branches (2 to 4,
 * FHTGS_OPT_VIEW_BRANCHES); the choices stay in the handle. Only parameters
 * that leave every image unchanged are tuned: K changes which fragments
 * enter the core, and the 16-px tiles define the foveation classes and the
 * sample grid. A 32x8 work tile would span two 16-px tiles whose instance
 * lists differ, so it would share no staging and is not offered. Leaves the
 * batch's images rendered. */
int fhtgs_autotune(fhtgs_handle h, const fhtgs_scene* scenes, int32_t n_scenes,
                   const fhtgs_view* views, const int32_t* view_scene, const int32_t* view_slot,
                   int32_t n_views, int32_t frames, void* stream);

/* The digit width the handle's sort uses for an image size (tuned or the
 * FHTGS_OPT_SORT_DIGIT_BITS default). */
int fhtgs_tuned_digit_bits(fhtgs_handle h, int32_t width, int32_t height, int32_t* bits);

/* The blend work tile the handle uses for an image size, as waves per
 * workgroup: 1 (8x8), 2 (16x8) or 4 (16x16) (tuned or the
 * FHTGS_OPT_BLEND_WORK_TILE default). */
int fhtgs_tuned_work_tile(fhtgs_handle h, int32_t width, int32_t height, int32_t* waves);

/* Device bytes one view slot needs, before any render (no handle, no
 * device call): *fixed_bytes for the per-Gaussian and per-tile buffers of
 * `view` over a scene of n_primitives Gaussians (with device-built culling
 * tables when the view brings no summed-area tables), *instance_bytes for
 * n_instances tile instances (keys, values and the sort's counters, under
 * the largest plan any digit width gives). A handle holds one slot per view
 * it renders -- per eye for stereo -- which only grows, plus a few KB of
 * parameters. The instance count of a frame is known after its preprocess;
 * frames rendered without a mid-frame wait (device-count and graph modes)
 * size their slot for the last exact count + 25 % + 65536. Reads only the
 * view's sizes, foveation radii and whether it brings summed-area tables
 * (its pointers may be NULL), and checks those as fhtgs_validate_view does
 * (the reason, on failure, from fhtgs_last_error(NULL)). The reference computes the same kind of figure
 * per call from element counts (required<T>, utils/buffer_utils.h:17-22) and
 * allocates it inside every call. */
int fhtgs_workspace_query(const fhtgs_view* view, int64_t n_primitives, int64_t n_instances,
                          int64_t* fixed_bytes, int64_t* instance_bytes);

/* Device scratch the handle currently holds (it grows on demand). */
int fhtgs_workspace_bytes(fhtgs_handle h, int64_t* bytes);

/* Stable LSD radix sort of (key, value) pairs in place on bits [0, end_bit);
 * key_bytes 2 or 4, values uint32. Uses the handle's scratch. */
int fhtgs_sort_pairs(fhtgs_handle h, void* keys, uint32_t* values, int64_t n,
                     int32_t key_bytes, int32_t end_bit, void* stream);

/* No reference counterpart. For callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * allocation): the stable LSD radix sort of fhtgs_sort_pairs (pairs sorted
 * in place on bits [0, end_bit), key_bytes 2 or 4, uint32 values; 8-bit
 * digits) with every piece of scratch -- alternate keys and values, digit
 * histograms, tickets, look-back status -- in one fixed device array of the
 * library (about 8.7 MB on each device the library runs on, held until the
 * process ends). It needs no handle, allocates nothing, and returns once the
 * work is enqueued on `stream` (NULL = the default stream) of the current
 * device; n above FHTGS_SORT_NOALLOC_MAX_PAIRS is rejected with
 * FHTGS_E_INVALID. `error` (device uint32, may be NULL) receives 0, or a
 * nonzero value if the sort detects an inconsistency. All calls on one
 * device share that device's array: calls that overlap on one device (two
 * streams, two threads) race. Failures are described by fhtgs_last_error(NULL). */
#define FHTGS_SORT_NOALLOC_MAX_PAIRS (1 << 20)
int fhtgs_sort_pairs_noalloc(void* keys, uint32_t* values, int64_t n, int32_t key_bytes,
                             int32_t end_bit, uint32_t* error, void* stream);

/* Culling tables on the device, in the layouts the render reads (the
 * reference takes them from its caller, torch_bindings/rasterization.py:28-30):
 * mask_out, ceil(gw*gh/32) words, bit t%32 of word t/32 for 16-px tile t,
 * from `visible` (gw*gh bytes, non-zero = visible; NULL = all visible);
 * vis_sat_out, (gw+1)*(gh+1) entries, x+(gw+1)*y = visible tiles in
 * [0,x)x[0,y) (needs mask_out); fovea_sat_out, (2R+1)^2 entries for R of
 * fhtgs_fovea_radius_tiles. Any output may be NULL; enqueued on `stream`.
 * A view whose two SAT pointers are both NULL gets them built this way inside
 * its frame (and, with a NULL mask, an all-visible mask). */
int fhtgs_build_tables(fhtgs_handle h, const uint8_t* visible, int32_t width, int32_t height,
                       int32_t foveation_radius_px, int32_t blend_width_px, uint32_t* mask_out,
                       uint32_t* vis_sat_out, uint32_t* fovea_sat_out, void* stream);

/* Draws the gaze marker into `image` (device float32, 3*H*W, layout by
 * to_chw): the gaze clamped to the image and truncated to a pixel, a marker of
 * `size` pixels -- type 0 circle, 1 square, 2 cross (3 px wide arms); rgb is
 * a host pointer to 3 floats. Enqueued on `stream`. */
int fhtgs_visualize_gaze(fhtgs_handle h, float* image, int32_t width, int32_t height,
                         float gaze_x, float gaze_y, const float* rgb, int32_t type, int32_t size,
                         int32_t to_chw, void* stream);

/* Draws the borders of the active tiles of the rasterizer's classification for
 * this gaze, mask and foveation radii: FOVEA sub-tiles orange, BLENDED
 * sub-tiles light blue, PERIPHERY tiles green. Returns after the overlay is drawn. */
int fhtgs_visualize_tile_boundaries(fhtgs_handle h, float* image, int32_t width, int32_t height,
                                    const uint32_t* visibility_mask, float gaze_x, float gaze_y,
                                    int32_t foveation_radius_px, int32_t blend_width_px,
                                    int32_t to_chw, void* stream);

/* Fovea and blend-ring radii in 16-px tiles for an image; the fovea table must
 * be built for the first. foveation_radius_px < 0 means foveation off. */
int fhtgs_fovea_radius_tiles(int32_t width, int32_t height, int32_t foveation_radius_px,
                             int32_t blend_width_px, int32_t* fovea_tiles, int32_t* blend_tiles);

/* Stage outputs of the last fhtgs_render, copied to host memory (synchronous). */
#define FHTGS_DBG_TILE_NUM 1        /* u32[n_large] */
#define FHTGS_DBG_TILE_MAP 2        /* u32[n_active] */
#define FHTGS_DBG_TILE_TYPE 3       /* u8 [n_active] */
#define FHTGS_DBG_TILE_MAP_PART 4   /* u32[n_active] */
#define FHTGS_DBG_PARTITION 6       /* u32[6] */
#define FHTGS_DBG_N_TOUCHED 7       /* u32[N] */
#define FHTGS_DBG_BOUNDS 8          /* u16[4N] */
#define FHTGS_DBG_VPMT1 9           /* f32[4N] */
#define FHTGS_DBG_VPMT2 10
#define FHTGS_DBG_VPMT4 11
#define FHTGS_DBG_RGBA 12
#define FHTGS_DBG_KEYS_UNSORTED 14  /* u32[I] (keys widened to 32 bits) */
#define FHTGS_DBG_VALS_UNSORTED 15  /* u32[I] */
#define FHTGS_DBG_KEYS_SORTED 16
#define FHTGS_DBG_VALS_SORTED 17
#define FHTGS_DBG_RANGES 18         /* u32[2*n_small] */
#define FHTGS_DBG_PARTITION_QUARTER 19 /* u32[2]: the quarter tiles' partition range */
/* u32[units]: the blend's work order. A unit is one active FOVEA, PERIPHERY or
 * BLENDED tile (its partition index) or a group of four QUARTER tiles (the
 * non-QUARTER count plus the group index); units come longest instance list
 * first, in cost buckets of a quarter octave. Scheduling only: no pixel
 * depends on it. */
#define FHTGS_DBG_BLEND_ORDER 20
int fhtgs_debug_read(fhtgs_handle h, int32_t view, int32_t id, void* host_dst,
                     int64_t dst_bytes, int64_t* bytes_written);

/* counts[16]: N, n_large, n_small, n_active, I, n_valid, end_bit, key_bytes,
 * n_visible(-1: not counted), gw, gh, fovea_tiles, blend_tiles, gaze_tx, gaze_ty, passes */
int fhtgs_debug_counts(fhtgs_handle h, int32_t view, int64_t* counts);

/* Keep a copy of the unsorted instance arrays for fhtgs_debug_read (costs
 * one extra copy per frame). */
#define FHTGS_OPT_KEEP_UNSORTED 1
/* Maximum radix digit width of the instance sort (8..11, default 11). */
#define FHTGS_OPT_SORT_DIGIT_BITS 2
/* Record per-stage HIP events; fhtgs_stage_times returns the last frame's ms. */
#define FHTGS_OPT_STAGE_TIMING 3
/* Blend tile-to-workgroup mapping: runs of this many consecutive tiles go to
 * one XCD (its own L2) so neighbouring tiles reuse primitive records; 0 or 1
 * keeps the launch order. Values 0..4096, default 16. Images do not depend on it. */
#define FHTGS_OPT_BLEND_XCD_RUN 4
/* How a frame waits for the instance count that sizes its sort:
 *   0  exact: read the counts back after the preprocess (one mid-frame wait),
 *      size the instance buffers to them; the reference's behaviour
 *   1  device counts: no mid-frame wait; kernels read the counts on the GPU
 *      and the buffers hold the last exact frame's count plus headroom
 *   2  as 1, with the part after the preprocess (instances to upsample)
 *      replayed from a captured hipGraph while the launch shape (views,
 *      sizes, K, scenes' sizes, buffers) stays the same (default)
 * A frame that needs more instances than its buffers hold in mode 1/2 is
 * rendered again in mode 0; the first frame of a new shape is always exact. */
#define FHTGS_OPT_FRAME_MODE 5
/* Concurrent view branches of an untimed frame: after the shared preprocess,
 * each view's chain (instances, sort, ranges, blend, upsample) runs on one of
 * this many streams of the handle (views round-robin), and tile
 * classification overlaps the preprocess; 1 = one stream. Values 1..4,
 * default 2. Images do not depend on it. */
#define FHTGS_OPT_VIEW_BRANCHES 6
/* Blend work tile, as waves per workgroup: 1 = 8x8 (one wave per 8x8 tile,
 * default), 2 = 16x8, 4 = 16x16: the sub-tiles of a FOVEA or BLENDED 16-px
 * tile share one staging of their (identical) instance list, and PERIPHERY
 * and QUARTER tiles go 2 or 4 waves to a workgroup. Setting it clears the
 * autotuned choices. Images do not depend on it. */
#define FHTGS_OPT_BLEND_WORK_TILE 7
int fhtgs_set_option(fhtgs_handle h, int32_t option, int64_t value);
int fhtgs_get_option(fhtgs_handle h, int32_t option, int64_t* value);

/* ms per stage of the last timed frame, all views of the call together:
 * clear, tiles, preprocess, scan (+ count readback), sync (host wait for the
 * counts), instances, sort, ranges, blend (+ output clear), upsample, total. */
#define FHTGS_NUM_STAGES 11
int fhtgs_stage_times(fhtgs_handle h, float* ms);

#ifdef __cplusplus
}
#endif

#endif /* FHTGS_H */
