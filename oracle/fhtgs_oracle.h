/*
 * CPU oracle for the foveated HTGS rasterizer: TEST INFRASTRUCTURE ONLY.
 *
 * A single-threaded, scalar restatement of the reference's inference path
 * (FoveatedHTGSCudaBackend/rasterization), used by the test suite as the
 * parity checker and by bench.py as the reported CPU baseline. It is never
 * linked into, loaded by, or called from the product library.
 *
 * Floating-point convention (shared with the GPU kernels, which are written
 * independently): IEEE single precision, no implicit contraction, and every
 * "a*b + c" of the reference expressions evaluated as one fmaf, with the
 * first product fused when two products are added ("a*b - c*d" is
 * fmaf(a, b, -(c*d))); dot products are fmaf chains left to right. expf/logf
 * and the other transcendentals come from fhtgs_math.h.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    const float* positions;   /* (N,3) */
    const float* scales;      /* (N,3) linear sigma */
    const float* rotations;   /* (N,4) quaternion (w,x,y,z), not normalised */
    const float* opacities;   /* (N) activated */
    const float* sh_0;        /* (N,1,3) */
    const float* sh_rest;     /* (N,S,3) */
    int32_t n_primitives;
    int32_t total_sh_bases;   /* S */
} orc_scene;

typedef struct {
    float M[12];              /* rows 0..2 of the world->view matrix, row major */
    float VPM[16];            /* world->screen, row major */
    float VPR_inv[16];        /* screen->world rotation part, row major */
    float cam_position[3];
    float gaze_position[2];   /* pixels, unclamped */
    const uint32_t* visibility_mask;
    const uint32_t* visibility_mask_area_table;
    const uint32_t* fovea_mask_area_table;
    const float* background_model_data;
    int32_t background_model_type;   /* 0 none, 1 SH, 2 texture, 3 raw texture */
    int32_t background_width;        /* texture (3,h,w): w */
    int32_t background_height;       /* texture: h */
    int32_t K;
    int32_t active_sh_bases;
    int32_t width;
    int32_t height;
    float focal_x, focal_y, center_x, center_y;
    float near_plane, far_plane, scale_modifier;
    int32_t to_chw;
    int32_t blur_periphery;
    int32_t anti_aliasing;
    int32_t foveation_radius_px;     /* reference: 375; <0 means foveation off */
    int32_t blend_width_px;          /* reference: 125 */
    int32_t blend_tile_stride;       /* >1: blend/upsample only every n-th active tile (timing samples) */
    int32_t quarter_radius_px;       /* extension: quarter-resolution level beyond this radius; <=0 off */
} orc_view;

/* Result arrays, fetched by id after orc_run. */
enum {
    ORC_COUNTS = 0,          /* int64[20], see ORC_C_* */
    ORC_TILE_NUM = 1,        /* u32[n_large]  0/1/4 per large tile */
    ORC_TILE_MAP = 2,        /* u32[n_active] compacted small-tile ids */
    ORC_TILE_TYPE = 3,       /* u8 [n_active] */
    ORC_TILE_MAP_PART = 4,   /* u32[n_active] after the stable type partition */
    ORC_TILE_TYPE_PART = 5,  /* u8 [n_active] */
    ORC_PARTITION = 6,       /* u32[8] (x,y) per type FOVEA, PERIPHERY, BLENDED, QUARTER */
    ORC_N_TOUCHED = 7,       /* u32[N] */
    ORC_BOUNDS = 8,          /* u16[4N] (xmin,xmax,ymin,ymax), valid where n_touched>0 */
    ORC_VPMT1 = 9,           /* f32[4N] */
    ORC_VPMT2 = 10,
    ORC_VPMT4 = 11,
    ORC_RGBA = 12,
    ORC_OFFSETS = 13,        /* u32[N] inclusive scan of n_touched */
    ORC_KEYS_UNSORTED = 14,  /* u32[I] (sentinel 0xffff or 0xffffffff) */
    ORC_VALS_UNSORTED = 15,  /* u32[I] (0xffffffff in sentinel slots) */
    ORC_KEYS_SORTED = 16,
    ORC_VALS_SORTED = 17,
    ORC_RANGES = 18,         /* u32[2*n_small] */
    ORC_IMAGE = 19,          /* f32[3HW] */
    ORC_IMAGE_TEMP = 20,     /* f32[3HW] (blur only) */
    ORC_TIMINGS = 21,        /* f64[8] ms: tiles, preprocess, scan, instances, sort, ranges, blend, upsample */
    ORC_NUM_ARRAYS = 22
};

enum {
    ORC_C_N = 0, ORC_C_N_LARGE = 1, ORC_C_N_SMALL = 2, ORC_C_N_ACTIVE = 3,
    ORC_C_INSTANCES = 4, ORC_C_VALID_INSTANCES = 5, ORC_C_END_BIT = 6,
    ORC_C_KEY_BYTES = 7, ORC_C_N_VISIBLE = 8, ORC_C_GRID_W = 9, ORC_C_GRID_H = 10,
    ORC_C_FOVEA_R_TILES = 11, ORC_C_BLEND_R_TILES = 12, ORC_C_GAZE_TX = 13,
    ORC_C_GAZE_TY = 14, ORC_C_EVALUATIONS = 15, ORC_C_PASS_CUTOFF = 16, ORC_C_PASS_ALPHA = 17,
    ORC_C_CORE_INSERTS = 18, ORC_C_QUARTER_R_TILES = 19
};

typedef struct orc_result orc_result;

/* Runs the whole pipeline; stop_after: 0 = everything, 1 = after preprocess,
 * 2 = after the sort. Returns NULL and fills err on invalid arguments. */
orc_result* orc_run(const orc_scene* scene, const orc_view* view, int32_t stop_after,
                    char* err, int32_t err_len);
/* The upsampling stage alone, on caller images (a test hook for the
 * interpolation and blur weights): the view's mask, gaze, radii and flags
 * classify the tiles as orc_run does, then PERIPHERY, BLENDED and QUARTER
 * tiles are filled in place in `image` (f32[3HW], layout by to_chw) or, with
 * blur_periphery, blurred from `image_temp` into `image`. Returns 0, or -1
 * with err filled. Tables other than the mask are not read. */
int32_t orc_upsample(const orc_view* view, float* image, float* image_temp, char* err, int32_t err_len);
int64_t orc_result_bytes(const orc_result* r, int32_t id);
int32_t orc_result_copy(const orc_result* r, int32_t id, void* dst, int64_t dst_bytes);
void orc_result_free(orc_result* r);

/* Table builder restated from the table layout the reference's preprocess reads
 * (kernel_utils.cuh:180-210): visibility SAT of (gw+1)*(gh+1) entries indexed
 * x + (gw+1)*y, fovea SAT of (2R+1)^2 entries of 3*[inside fovea circle]. */
void orc_build_visibility_sat(const uint32_t* mask, int32_t gw, int32_t gh, uint32_t* sat);
void orc_build_fovea_sat(int32_t radius_tiles, uint32_t* sat);

/* Debug overlays (visualization/src/visualization.cu): the gaze marker (type
 * 0 circle, 1 square, 2 cross) and the tile-type borders of the render's
 * classification, drawn into a float32 image (HWC, or CHW when chw). */
void orc_visualize_gaze(float* image, int32_t width, int32_t height, float gaze_x, float gaze_y,
                        const float* rgb, int32_t type, int32_t size, int32_t chw);
void orc_visualize_tile_boundaries(float* image, int32_t width, int32_t height, const uint32_t* mask,
                                   float gaze_x, float gaze_y, int32_t foveation_radius_px,
                                   int32_t blend_width_px, int32_t chw);

/* Stable LSD reference sort: keys compared on bits [0, end_bit). */
void orc_stable_sort_pairs(uint32_t* keys, uint32_t* vals, int64_t n, int32_t end_bit);

/* Scalar math probes used by the self tests. */
float orc_expf(float x);
float orc_logf(float x);
float orc_atan2f(float y, float x);
float orc_asinf(float x);
float orc_tanf(float x);
float orc_tanhf(float x);
/* SH colour (degree by active bases) for a unit direction; sh_rest holds 15 x 3. */
void orc_sh_eval(const float* sh0, const float* sh_rest, const float* dir, int32_t active, float* rgb);

#ifdef __cplusplus
}
#endif
