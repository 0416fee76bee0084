// Host pipeline and C ABI of libfhtgs.so (declared in include/fhtgs.h).
//
// Per frame and view (rasterization/src/inference.cu:49-472 is the reference
// driver): clear -> tile classification -> preprocess (+ block sums of
// tiles-touched) -> scan of the block sums; one wait on the stream for the
// instance and tile counts; then sort clear -> instances (+ digit histograms)
// -> onesweep passes -> ranges -> output clear -> blend -> upsample.
// Scratch lives in the handle and only grows, so a steady stream of frames
// allocates nothing.
#include "../../include/fhtgs.h"
#include "fhtgs_internal.h"

#include <algorithm>
#include <cmath>
#include <mutex>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

using namespace fhtgs;

namespace {

constexpr size_t kAlign = 256;

size_t align_up(size_t x) { return (x + kAlign - 1) & ~(kAlign - 1); }

int end_bit_of(uint32_t n) {
    int b = 0;
    while (n) { ++b; n >>= 1; }
    return b;
}

void fov_radii(int width, int height, int fr_px, int bw_px, int& R, int& B) {
    const int gw = (width + kTileL - 1) / kTileL, gh = (height + kTileL - 1) / kTileL;
    if (fr_px < 0) {
        const int cap = (int)std::ceil(std::sqrt((double)gw * gw + (double)gh * gh)) + 1;
        R = cap;
        B = cap;
    } else {
        R = (fr_px + kTileL - 1) / kTileL;
        const int raw = fr_px - bw_px;
        B = raw > 0 ? raw / kTileL : 0;
    }
}

SortPlan make_plan(int end_bit, int max_bits) {
    SortPlan p;
    std::memset(&p, 0, sizeof(p));
    p.end_bit = end_bit;
    p.key_bytes = end_bit <= 16 ? 2 : 4;
    max_bits = std::max(4, std::min(11, max_bits));
    p.passes = std::max(1, (end_bit + max_bits - 1) / max_bits);
    const int per = (end_bit + p.passes - 1) / p.passes;
    int shift = 0;
    p.bins_max = 0;
    for (int i = 0; i < p.passes; ++i) {
        const int b = std::min(per, end_bit - shift);
        p.bits[i] = std::max(1, b);
        p.shift[i] = shift;
        shift += b;
        p.bins_max = std::max(p.bins_max, 1 << p.bits[i]);
    }
    return p;
}

struct DeviceBuffer {
    void* ptr = nullptr;
    size_t cap = 0;
};

struct ViewSlot {
    DeviceBuffer fixed;
    DeviceBuffer inst;
    DeviceBuffer debug;
    uint32_t* host_fi = nullptr;  // pinned, FI_WORDS
    ViewParams vp;                // host copy
    ViewParams* dvp = nullptr;    // device copy, in the handle's parameter block
    ViewBuffers vb;
    InstanceBuffers ib;
    SortPlan plan;
    int64_t counts[16];
    int sorted_src = 0;
    uint32_t* dbg_keys = nullptr;
    uint32_t* dbg_vals = nullptr;
    uint32_t n_instances = 0;
    uint32_t n_active = 0;
    int64_t n_prims = 0;
    // instance capacity for frames that do not read the counts back, from
    // the last exact frame of the same shape (N, n_small); 0 = unknown
    int64_t cap_hint = 0;
    int64_t cap_for_n = -1, cap_for_tiles = -1;
    bool auto_tables = false;  // culling tables built on the device each frame
    bool own_mask = false;     // ... including the mask (all tiles visible)
};

// One preprocess launch: a scene and one or two of its views.
// Views whose Gaussians one preprocess pass reads once: one scene, equal
// preprocess parameters, at most kMaxGroup views.
constexpr int kMaxGroup = 4;
struct PreGroup {
    int scene;
    int nv;
    int v[kMaxGroup];
};

enum FrameMode { kModeExact = 0, kModeDevice = 1, kModeGraph = 2 };

struct GraphEntry {
    std::vector<int64_t> sig;
    hipGraphExec_t exec = nullptr;
    uint64_t last_use = 0;
};
constexpr size_t kMaxGraphs = 64;
constexpr int kMaxSlots = 4096;
constexpr int kMaxBranches = 4;

}  // namespace

struct fhtgs_context {
    int device = 0;
    std::string err;
    std::vector<ViewSlot> slots;
    DeviceBuffer sort_scratch;
    hipEvent_t done = nullptr;
    bool have_done = false;
    int keep_unsorted = 0;
    int digit_bits = 11;
    int xcd_run = 16;
    int work_tile = 1;  // waves per blend workgroup: 1 (8x8), 2 (16x8) or 4 (16x16 work tiles)
    int timing = 0;
    int frame_mode = kModeGraph;
    int last_mode = kModeExact;
    // concurrent view branches: the frame's stream plus side streams (default
    // priority), forked and joined with events
    int branches = 2;
    hipStream_t side[kMaxBranches - 1] = {};
    hipEvent_t ev_fork = nullptr;
    hipEvent_t ev_params = nullptr;  // the parameter upload, on side stream 0
    hipEvent_t ev_join[kMaxBranches - 1] = {};
    hipEvent_t ev[FHTGS_NUM_STAGES + 1] = {};
    float stage_ms[FHTGS_NUM_STAGES] = {};
    // per-frame parameters: ViewParams of every view, then SceneArgs of every
    // scene; written to pinned staging and copied up once per frame
    DeviceBuffer params;
    void* params_host = nullptr;
    size_t params_host_cap = 0;
    // the slots of the views of the call in progress (view v -> slot)
    std::vector<ViewSlot*> cur;
    // graph replay of device-count frames: captured graphs by launch
    // signature, least recently used evicted beyond kMaxGraphs
    hipStream_t cap_stream = nullptr;
    std::vector<GraphEntry> graphs;
    uint64_t graph_clock = 0;
    // autotuned sort digit width per view size (n_small); others use digit_bits
    std::vector<std::pair<int64_t, int>> tuned_bits;
    // autotuned blend work tile per view size (n_small); others use work_tile
    std::vector<std::pair<int64_t, int>> tuned_wt;
    // tile-boundary overlay: its own parameters and tile buffers
    ViewSlot vis;
    DeviceBuffer vis_params;
    ViewParams* vis_params_host = nullptr;
};

// The blend work tile of a view size (autotuned, or the handle's default).
static int work_tile_for(const fhtgs_context* h, int64_t n_small) {
    for (const auto& t : h->tuned_wt)
        if (t.first == n_small) return t.second;
    return h->work_tile;
}

namespace {

// Graph capture on one handle's stream and allocation or freeing of device
// or pinned memory by another handle must not overlap (the runtime refuses
// such calls, or waits on the capturing stream): both hold this lock.
std::mutex& device_memory_lock() {
    static std::mutex m;
    return m;
}

// the message of the last failure of a call without a handle, per thread
thread_local std::string g_no_handle_err;

const char* const kTablesMismatch =
    "the culling tables do not fit the view (a summed-area table of another image size or "
    "foveation radius): tiles-touched exceeds 4 per tile of a Gaussian's bounds";

int set_err(fhtgs_context* h, int code, const std::string& msg) {
    if (h) h->err = msg;
    else g_no_handle_err = msg;
    return code;
}

#define HCHK(expr)                                                                          \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return set_err(h, FHTGS_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

int ensure(fhtgs_context* h, DeviceBuffer& b, size_t bytes) {
    if (bytes <= b.cap) return FHTGS_OK;
    std::lock_guard<std::mutex> lock(device_memory_lock());
    size_t want = std::max(bytes, b.cap + b.cap / 4);
    want = align_up(want);
    if (b.ptr) {
        // the handle's last frame may still use the buffer; other handles'
        // work is none of its business (a device-wide wait would also fail
        // while another thread captures a graph)
        if (h->have_done) HCHK(hipEventSynchronize(h->done));
        HCHK(hipFree(b.ptr));
        b.ptr = nullptr;
        b.cap = 0;
    }
    hipError_t e = hipMalloc(&b.ptr, want);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        b.ptr = nullptr;
        return set_err(h, FHTGS_E_OUT_OF_MEMORY,
                       "device scratch allocation of " + std::to_string(want) + " bytes failed");
    }
    b.cap = want;
    return FHTGS_OK;
}

struct Carver {
    char* base;
    size_t off = 0;
    explicit Carver(void* b) : base((char*)b) {}
    template <typename T>
    T* take(size_t count) {
        off = align_up(off);
        T* p = (T*)(base ? base + off : nullptr);
        off += sizeof(T) * count;
        return p;
    }
};

void carve_fixed(void* base, int64_t N, int n_large, int n_small, ViewBuffers& vb, size_t& bytes,
                 int gw = 0, int gh = 0, int R = -1) {
    Carver c(base);
    const size_t nb = (size_t)((N + kPrimBlock - 1) / kPrimBlock);
    vb.frame_info = c.take<uint32_t>(FI_WORDS);
    vb.tile_num = c.take<uint32_t>((size_t)n_large);
    vb.tile_map = c.take<uint32_t>((size_t)n_small);
    vb.tile_type = c.take<uint8_t>((size_t)n_small);
    vb.tile_map_part = c.take<uint32_t>((size_t)n_small);
    vb.ranges = c.take<uint2>((size_t)n_small);
    vb.blend_order = c.take<uint32_t>((size_t)n_small);
    vb.order_bucket = c.take<uint8_t>((size_t)n_small);
    vb.order_counts = c.take<uint32_t>(128);
    vb.n_touched = c.take<uint32_t>((size_t)N);
    vb.bounds = c.take<ushort4>((size_t)N);
    vb.prim = c.take<float4>((size_t)N * 4);
    vb.ell = c.take<float4>((size_t)N * 2);
    vb.block_sums = c.take<uint32_t>(std::max<size_t>(nb, 1));
    vb.block_vis = c.take<uint32_t>(std::max<size_t>(nb, 1));
    vb.tab_mask = vb.tab_vis_sat = vb.tab_fovea_sat = nullptr;
    if (R >= 0) {
        vb.tab_mask = c.take<uint32_t>((size_t)(n_large + 31) / 32);
        vb.tab_vis_sat = c.take<uint32_t>((size_t)(gw + 1) * (gh + 1));
        vb.tab_fovea_sat = c.take<uint32_t>((size_t)(2 * R + 1) * (2 * R + 1));
    }
    bytes = align_up(c.off);
}

void carve_inst(void* base, int64_t I, const SortPlan& plan, InstanceBuffers& ib, size_t& bytes) {
    Carver c(base);
    const int64_t cap = std::max<int64_t>(I, 1);
    const int64_t tiles = (cap + kSortTile - 1) / kSortTile;
    ib.keys[0] = c.take<char>((size_t)cap * plan.key_bytes);
    ib.keys[1] = c.take<char>((size_t)cap * plan.key_bytes);
    ib.vals[0] = c.take<uint32_t>((size_t)cap);
    ib.vals[1] = c.take<uint32_t>((size_t)cap);
    ib.hist = c.take<uint32_t>((size_t)kMaxPasses * kMaxBins);
    ib.tickets = c.take<uint32_t>((size_t)kMaxPasses + 1);
    ib.status = c.take<uint32_t>((size_t)plan.passes * (size_t)tiles * plan.bins_max);
    ib.capacity = cap;
    ib.tiles_cap = tiles;
    bytes = align_up(c.off);
}

// The checks that depend on the view's sizes and radii only.
int validate_geometry(fhtgs_context* h, const fhtgs_view& v, const std::string& p) {
    if (v.width <= 0 || v.height <= 0) return set_err(h, FHTGS_E_INVALID, p + "width and height must be positive");
    if ((int64_t)v.width * v.height * 3 > (int64_t)1 << 31)
        return set_err(h, FHTGS_E_INVALID, p + "image too large");
    if (v.foveation_radius_px >= 0 && v.blend_width_px > v.foveation_radius_px)
        return set_err(h, FHTGS_E_INVALID, p + "blend width exceeds the foveation radius");
    // the ring's cross-fade divides by blend_width_px - 16 (interpolation.cuh:244-250)
    if (v.foveation_radius_px >= 0 && v.blend_width_px <= kTileL)
        return set_err(h, FHTGS_E_INVALID, p + "blend width must exceed 16 px with foveation on");
    const int64_t n_large = (int64_t)((v.width + kTileL - 1) / kTileL) * ((v.height + kTileL - 1) / kTileL);
    if (n_large > kMaxLargeTiles)
        return set_err(h, FHTGS_E_UNSUPPORTED, p + "more than 65536 16-px tiles (above 4096 x 4096 pixels)");
    if (v.quarter_radius_px > 0 && (v.foveation_radius_px < 0 || v.quarter_radius_px < v.foveation_radius_px))
        return set_err(h, FHTGS_E_INVALID, p + "the quarter radius needs foveation and must be >= its radius");
    return FHTGS_OK;
}

int validate_view(fhtgs_context* h, const fhtgs_view& v, int idx) {
    const std::string p = "view " + std::to_string(idx) + ": ";
    if (const int rc = validate_geometry(h, v, p)) return rc;
    if (!v.M || !v.VPM || !v.cam_position || !v.image)
        return set_err(h, FHTGS_E_INVALID, p + "null camera or image pointer");
    // both tables given, or neither (built on the device each frame)
    if ((v.visibility_mask_area_table == nullptr) != (v.fovea_mask_area_table == nullptr))
        return set_err(h, FHTGS_E_INVALID, p + "give both summed-area tables or neither");
    if (v.visibility_mask_area_table && !v.visibility_mask)
        return set_err(h, FHTGS_E_INVALID, p + "null visibility mask");
    if (v.blur_periphery && !v.image_temp)
        return set_err(h, FHTGS_E_INVALID, p + "blur_periphery needs image_temp");
    if (v.background_model_type < 0 || v.background_model_type > 3)
        return set_err(h, FHTGS_E_INVALID, p + "unknown background model type");
    if (v.background_model_type != 0 && (!v.background_model_data || !v.VPR_inv))
        return set_err(h, FHTGS_E_INVALID, p + "background model needs data and VPR_inv");
    if ((v.background_model_type == 2 || v.background_model_type == 3) &&
        (v.background_width <= 0 || v.background_height <= 0))
        return set_err(h, FHTGS_E_INVALID, p + "texture background needs its size");
    if (v.quarter_radius_px > 0 && v.blur_periphery)
        return set_err(h, FHTGS_E_UNSUPPORTED, p + "the quarter level does not combine with blur_periphery");
    return FHTGS_OK;
}

void fill_params(const fhtgs_scene& sc, const fhtgs_view& v, ViewParams& vp) {
    std::memset(&vp, 0, sizeof(vp));
    vp.M = v.M;
    vp.VPM = v.VPM;
    vp.VPR_inv = v.VPR_inv;
    vp.cam = v.cam_position;
    vp.vis_mask = v.visibility_mask;
    vp.vis_sat = v.visibility_mask_area_table;
    vp.fovea_sat = v.fovea_mask_area_table;
    vp.bg_data = v.background_model_data;
    vp.bg_type = v.background_model_type;
    vp.bg_w = v.background_width;
    vp.bg_h = v.background_height;
    vp.width = v.width;
    vp.height = v.height;
    vp.gw = (v.width + kTileL - 1) / kTileL;
    vp.gh = (v.height + kTileL - 1) / kTileL;
    vp.n_large = vp.gw * vp.gh;
    vp.n_small = vp.n_large * kNumSub;
    fov_radii(v.width, v.height, v.foveation_radius_px, v.blend_width_px, vp.fovea_r, vp.blend_r);
    vp.blend_radius_px = vp.blend_r * kTileL;
    vp.blend_width_px = v.blend_width_px;
    vp.quarter_r = v.quarter_radius_px > 0 ? (v.quarter_radius_px + kTileL - 1) / kTileL : 0;
    // rasterization/src/inference.cu:79-105: clamp, then round to a large
    // tile; the clamp bound is the small-grid size, as in the reference
    vp.gcx = std::fmax(0.0f, std::fmin(v.gaze_x, (float)(v.width - 1)));
    vp.gcy = std::fmax(0.0f, std::fmin(v.gaze_y, (float)(v.height - 1)));
    const int gsw = vp.gw * kStride, gsh = vp.gh * kStride;
    vp.gtx = std::max(0, std::min(gsw - 1, ((int)vp.gcx + kTileL / 2) / kTileL));
    vp.gty = std::max(0, std::min(gsh - 1, ((int)vp.gcy + kTileL / 2) / kTileL));
    vp.active_sh = v.active_sh_bases;
    vp.total_sh = sc.total_sh_bases;
    vp.fx = v.focal_x;
    vp.fy = v.focal_y;
    vp.cx = v.center_x;
    vp.cy = v.center_y;
    vp.near_plane = v.near_plane;
    vp.far_plane = v.far_plane;
    vp.scale_mod = v.scale_modifier;
    vp.to_chw = v.to_chw ? 1 : 0;
    vp.blur = v.blur_periphery ? 1 : 0;
    vp.aaa = v.anti_aliasing ? 1 : 0;
    vp.image_final = v.image;
    vp.image_blend = v.blur_periphery ? v.image_temp : v.image;
}

int record(fhtgs_context* h, int stage, hipStream_t s) {
    if (!h->timing) return FHTGS_OK;
    HCHK(hipEventRecord(h->ev[stage], s));
    return FHTGS_OK;
}

bool same_preprocess(const fhtgs_view& a, const fhtgs_view& b) {
    return a.active_sh_bases == b.active_sh_bases && a.near_plane == b.near_plane &&
           a.far_plane == b.far_plane && a.scale_modifier == b.scale_modifier &&
           a.anti_aliasing == b.anti_aliasing;
}

int64_t with_headroom(int64_t I) {
    const int64_t c = I + I / 4 + 65536;
    return (c + kSortTile - 1) / kSortTile * kSortTile;
}

// Everything the host decides when it enqueues a device-count frame: the
// same signature launches the same kernels on the same buffers, so a
// captured graph can be replayed with only the parameter block rewritten.
std::vector<int64_t> launch_signature(const fhtgs_context* h, const std::vector<PreGroup>& groups,
                                      const fhtgs_scene* scenes, const fhtgs_view* views, int n_views) {
    std::vector<int64_t> g;
    g.push_back(n_views);
    g.push_back(h->digit_bits);
    g.push_back(h->xcd_run);
    g.push_back(h->branches);
    g.push_back((int64_t)(uintptr_t)h->params.ptr);
    for (const PreGroup& pg : groups) {
        g.push_back(pg.scene);
        g.push_back(pg.nv);
        for (int k = 0; k < kMaxGroup; ++k) g.push_back(k < pg.nv ? pg.v[k] : -1);
        g.push_back(scenes[pg.scene].n_primitives);
        g.push_back(scenes[pg.scene].total_sh_bases);
        g.push_back(views[pg.v[0]].anti_aliasing ? 1 : 0);
    }
    for (int v = 0; v < n_views; ++v) {
        const ViewSlot& sl = (*h->cur[v]);
        g.push_back(sl.vp.width);
        g.push_back(sl.vp.height);
        g.push_back(views[v].K);
        g.push_back(views[v].clear_output ? 1 : 0);
        g.push_back(sl.auto_tables ? 1 + (sl.own_mask ? 1 : 0) : 0);
        g.push_back(sl.auto_tables ? sl.vp.fovea_r : -1);
        g.push_back(sl.ib.capacity);
        g.push_back(sl.plan.passes);
        g.push_back(sl.plan.bits[0]);
        g.push_back(work_tile_for(h, sl.vp.n_small));
        g.push_back((int64_t)(uintptr_t)sl.fixed.ptr);
        g.push_back((int64_t)(uintptr_t)sl.inst.ptr);
    }
    return g;
}

// Branch b of a frame on stream s: s itself for b = 0, else a side stream.
hipStream_t branch_stream(fhtgs_context* h, int b, hipStream_t s) { return b == 0 ? s : h->side[b - 1]; }

// Side streams 1..nb-1 wait for the work enqueued on s so far.
int fork_branches(fhtgs_context* h, int nb, hipStream_t s) {
    if (nb <= 1) return FHTGS_OK;
    HCHK(hipEventRecord(h->ev_fork, s));
    for (int b = 1; b < nb; ++b) HCHK(hipStreamWaitEvent(h->side[b - 1], h->ev_fork, 0));
    return FHTGS_OK;
}

// s waits for the work enqueued on side streams 1..nb-1.
int join_branches(fhtgs_context* h, int nb, hipStream_t s) {
    for (int b = 1; b < nb; ++b) {
        HCHK(hipEventRecord(h->ev_join[b - 1], h->side[b - 1]));
        HCHK(hipStreamWaitEvent(s, h->ev_join[b - 1], 0));
    }
    return FHTGS_OK;
}

// Phase 1 of a frame: parameter upload, clears, tile classification,
// preprocess, scan. pbytes: the parameter block's size. Untimed, with
// branches, the preprocess (which takes its parameters as kernel arguments
// and writes its counts per block) is the frame's first kernel, and the
// clears, the upload and the tile classifications (one workgroup each) run
// on the side streams beside it and are joined into the stream behind it -- they finish
// long before it does -- so the views' scans (one launch, one workgroup a
// view) follow the preprocess directly and phase 2 starts right after them,
// with no round trip through a side stream. sort_clear: the
// instance buffers are already sized (device counts, graph), so their sort
// counters are cleared here, after each view's tile classification, instead
// of in phase 2.
int enqueue_phase1(fhtgs_context* h, const std::vector<PreGroup>& groups, const fhtgs_scene* scenes,
                   size_t pbytes, int n_views, bool timed, bool sort_clear, hipStream_t s) {
    int rc = 0;
    const int nb = timed ? 1 : std::min(h->branches, n_views + 1);
    if (timed && (rc = record(h, 0, s))) return rc;
    std::vector<const ViewParams*> cvp((size_t)n_views);
    std::vector<const ViewBuffers*> cvb((size_t)n_views);
    for (int v = 0; v < n_views; ++v) {
        ViewSlot& sl = *h->cur[v];
        if (sl.auto_tables)
            HCHK(launch_build_tables(nullptr, sl.own_mask, sl.vp.gw, sl.vp.gh, sl.vp.fovea_r,
                                     const_cast<uint32_t*>(sl.vp.vis_mask), sl.vb.tab_vis_sat, sl.vb.tab_fovea_sat, s));
        cvp[(size_t)v] = &sl.vp;
        cvb[(size_t)v] = &sl.vb;
    }
    if (nb == 1) HCHK(launch_clear_frames(cvp.data(), cvb.data(), n_views, s));
    if (timed && (rc = record(h, 1, s))) return rc;
    auto tiles = [&]() -> int {
        for (int v = 0; v < n_views; ++v) {
            ViewSlot& sl = *h->cur[v];
            hipStream_t st = branch_stream(h, nb > 1 ? 1 + v % (nb - 1) : 0, s);
            HCHK(launch_tiles(sl.vp, sl.dvp, sl.vb, st));
            if (sort_clear) HCHK(launch_sort_clear(sl.ib, sl.plan, sl.ib.tiles_cap, st));
        }
        return FHTGS_OK;
    };
    auto preprocess = [&]() -> int {
        for (const PreGroup& pg : groups) {
            const ViewParams* hv[kMaxGroup];
            ViewBuffers vbs[kMaxGroup];
            for (int k = 0; k < pg.nv; ++k) {
                hv[k] = &(*h->cur[pg.v[k]]).vp;
                vbs[k] = (*h->cur[pg.v[k]]).vb;
            }
            const fhtgs_scene& sc = scenes[pg.scene];
            const SceneArgs sa{sc.positions, sc.scales, sc.rotations, sc.opacities, sc.sh_0, sc.sh_rest};
            HCHK(launch_preprocess(hv, vbs, pg.nv, sa, (uint32_t)sc.n_primitives, s));
        }
        return FHTGS_OK;
    };
    if (nb > 1) {
        // the preprocess is enqueued first; the clears, the upload and then
        // the tile kernels follow on the side streams
        HCHK(hipEventRecord(h->ev_fork, s));
        if ((rc = preprocess())) return rc;
        HCHK(hipStreamWaitEvent(h->side[0], h->ev_fork, 0));
        // the preprocess reads no frame-info word (it writes per-block counts),
        // so the clears run here, beside it, ahead of the tile kernels
        HCHK(launch_clear_frames(cvp.data(), cvb.data(), n_views, h->side[0]));
        HCHK(hipMemcpyAsync(h->params.ptr, h->params_host, pbytes, hipMemcpyHostToDevice, h->side[0]));
        if (nb > 2) {
            HCHK(hipEventRecord(h->ev_params, h->side[0]));
            for (int b = 2; b < nb; ++b) HCHK(hipStreamWaitEvent(h->side[b - 1], h->ev_params, 0));
        }
        if ((rc = tiles())) return rc;
        if ((rc = join_branches(h, nb, s))) return rc;
    } else {
        HCHK(hipMemcpyAsync(h->params.ptr, h->params_host, pbytes, hipMemcpyHostToDevice, s));
        if ((rc = tiles())) return rc;
        if (timed && (rc = record(h, 2, s))) return rc;
        if ((rc = preprocess())) return rc;
    }
    if (timed && (rc = record(h, 3, s))) return rc;
    // the views' block-sum scans
    std::vector<const ViewBuffers*> svb;
    std::vector<uint32_t> snb;
    for (int v = 0; v < n_views; ++v) {
        const int64_t N = (*h->cur[v]).n_prims;
        if (N > 0) {
            svb.push_back(&(*h->cur[v]).vb);
            snb.push_back((uint32_t)((N + kPrimBlock - 1) / kPrimBlock));
        }
    }
    if (!svb.empty()) HCHK(launch_scan_blocks(svb.data(), snb.data(), (int)svb.size(), s));
    return FHTGS_OK;
}

// Phase 2: instances, sort, ranges, output clear + blend, upsample.
// exact: grids from the counts read back; otherwise from the capacities.
int enqueue_instances(fhtgs_context* h, ViewSlot& slot, bool exact, hipStream_t s) {
    // (non-exact frames cleared the sort counters in phase 1)
    if (exact)
        HCHK(launch_sort_clear(slot.ib, slot.plan, ((int64_t)slot.n_instances + kSortTile - 1) / kSortTile, s));
    HCHK(launch_instances(slot.dvp, slot.vb, slot.ib, slot.plan, (uint32_t)slot.n_prims, slot.vp.n_large, s));
    if (exact && h->keep_unsorted && slot.n_instances > 0) {
        HCHK(launch_widen_keys(slot.ib.keys[0], slot.plan.key_bytes, slot.dbg_keys, slot.n_instances, s));
        HCHK(hipMemcpyAsync(slot.dbg_vals, slot.ib.vals[0], (size_t)slot.n_instances * 4, hipMemcpyDeviceToDevice,
                            s));
    }
    return FHTGS_OK;
}

int enqueue_sort(fhtgs_context* h, ViewSlot& slot, bool exact, hipStream_t s) {
    (void)h;
    const int64_t tiles = exact ? ((int64_t)slot.n_instances + kSortTile - 1) / kSortTile : slot.ib.tiles_cap;
    int src = 0;
    for (int p = 0; p < slot.plan.passes; ++p) {
        HCHK(launch_sort_pass(slot.ib, slot.plan, p, src, tiles, slot.vb.frame_info, 0u, s));
        src = 1 - src;
    }
    slot.sorted_src = src;
    return FHTGS_OK;
}

int enqueue_ranges(fhtgs_context* h, ViewSlot& slot, bool exact, hipStream_t s) {
    (void)h;
    HCHK(launch_ranges(slot.vb, slot.ib, slot.plan, slot.sorted_src,
                       exact ? slot.n_instances : (uint32_t)slot.ib.capacity, s));
    return FHTGS_OK;
}

int enqueue_blend(fhtgs_context* h, ViewSlot& slot, const fhtgs_view& view, bool exact, hipStream_t s) {
    if (view.clear_output) HCHK(launch_clear_image(slot.vp, slot.dvp, slot.vb, s));
    const int wt = work_tile_for(h, slot.vp.n_small);
    // a bound on the work units: n_small single-wave units, or n_small / W + 2
    uint32_t grid = (uint32_t)(wt <= 1 ? slot.vp.n_small : slot.vp.n_small / wt + 2);
    if (exact) grid = blend_units(slot.host_fi + FI_PART, slot.n_active, wt);
    HCHK(launch_blend_order(slot.vb, slot.vp.n_small, wt, s));
    HCHK(launch_blend(slot.dvp, slot.vb, slot.ib.vals[slot.sorted_src], view.K, grid, h->xcd_run, wt,
                      (uint32_t)slot.vp.n_small, s));
    return FHTGS_OK;
}

int enqueue_upsample(fhtgs_context* h, ViewSlot& slot, bool exact, hipStream_t s) {
    (void)h;
    uint32_t grid = (uint32_t)slot.vp.n_large;
    if (exact) {
        const uint32_t* part = slot.host_fi + FI_PART;
        grid = (part[3] - part[2]) + (part[5] - part[4] + 3u) / 4u + (part[7] - part[6]);
    }
    HCHK(launch_upsample(slot.dvp, slot.vb, grid, s));
    return FHTGS_OK;
}

// Timed: each stage for every view before the next, one event pair per
// stage. Untimed: every view's chain on its branch (views round-robin over
// the branches), so one view's sort overlaps another's blend.
int enqueue_phase2(fhtgs_context* h, const fhtgs_view* views, int n_views, bool exact, bool timed,
                   hipStream_t s) {
    int rc = 0;
    if (!timed) {
        const int nb = std::min(h->branches, n_views);
        if ((rc = fork_branches(h, nb, s))) return rc;
        for (int v = 0; v < n_views; ++v) {
            ViewSlot& slot = (*h->cur[v]);
            hipStream_t st = branch_stream(h, v % nb, s);
            if ((rc = enqueue_instances(h, slot, exact, st)) || (rc = enqueue_sort(h, slot, exact, st)) ||
                (rc = enqueue_ranges(h, slot, exact, st)) || (rc = enqueue_blend(h, slot, views[v], exact, st)) ||
                (rc = enqueue_upsample(h, slot, exact, st)))
                return rc;
        }
        return join_branches(h, nb, s);
    }
    if ((rc = record(h, 5, s))) return rc;
    for (int v = 0; v < n_views; ++v)
        if ((rc = enqueue_instances(h, *h->cur[v], exact, s))) return rc;
    if ((rc = record(h, 6, s))) return rc;
    for (int v = 0; v < n_views; ++v)
        if ((rc = enqueue_sort(h, *h->cur[v], exact, s))) return rc;
    if ((rc = record(h, 7, s))) return rc;
    for (int v = 0; v < n_views; ++v)
        if ((rc = enqueue_ranges(h, *h->cur[v], exact, s))) return rc;
    if ((rc = record(h, 8, s))) return rc;
    for (int v = 0; v < n_views; ++v)
        if ((rc = enqueue_blend(h, *h->cur[v], views[v], exact, s))) return rc;
    if ((rc = record(h, 9, s))) return rc;
    for (int v = 0; v < n_views; ++v)
        if ((rc = enqueue_upsample(h, *h->cur[v], exact, s))) return rc;
    if ((rc = record(h, 10, s))) return rc;
    return FHTGS_OK;
}

// Instance buffers of a view for `cap` instances.
int digit_bits_for(const fhtgs_context* h, int64_t n_small) {
    for (const auto& t : h->tuned_bits)
        if (t.first == n_small) return t.second;
    return h->digit_bits;
}


int size_instances(fhtgs_context* h, ViewSlot& slot, int64_t cap) {
    const int end_bit = end_bit_of((uint32_t)slot.vp.n_small + 1u);
    slot.plan = make_plan(end_bit, digit_bits_for(h, slot.vp.n_small));
    size_t bytes = 0;
    carve_inst(nullptr, cap, slot.plan, slot.ib, bytes);
    int rc = ensure(h, slot.inst, bytes);
    if (rc) return rc;
    carve_inst(slot.inst.ptr, cap, slot.plan, slot.ib, bytes);
    return FHTGS_OK;
}

int render_impl(fhtgs_context* h, const fhtgs_scene* scenes, int n_scenes, const fhtgs_view* views,
                const int32_t* view_scene, const int32_t* view_slot, int n_views, bool pair_views,
                hipStream_t s, int force_mode) {
    // ---- preprocess groups: up to kMaxGroup views of one scene that agree
    // on the preprocess parameters share one pass (each view's outputs depend
    // on that view alone), whether or not they are adjacent in the call: each
    // view is grouped with the next ungrouped ones that qualify
    std::vector<PreGroup> groups;
    std::vector<char> grouped((size_t)n_views, 0);
    for (int v = 0; v < n_views; ++v) {
        if (grouped[(size_t)v]) continue;
        PreGroup pg{view_scene[v], 1, {v, v, v, v}};
        grouped[(size_t)v] = 1;
        for (int w = v + 1; pair_views && w < n_views && pg.nv < kMaxGroup; ++w) {
            if (!grouped[(size_t)w] && view_scene[w] == view_scene[v] && same_preprocess(views[v], views[w])) {
                pg.v[pg.nv++] = w;
                grouped[(size_t)w] = 1;
            }
        }
        groups.push_back(pg);
    }
    // view v renders into slot view_slot[v] (default v): a slot keeps its
    // buffers and instance capacity from call to call
    int max_slot = n_views - 1;
    if (view_slot)
        for (int v = 0; v < n_views; ++v) max_slot = std::max(max_slot, (int)view_slot[v]);
    if ((int)h->slots.size() <= max_slot) h->slots.resize((size_t)max_slot + 1);
    h->cur.resize((size_t)n_views);
    for (int v = 0; v < n_views; ++v) h->cur[(size_t)v] = &h->slots[(size_t)(view_slot ? view_slot[v] : v)];

    // ---- parameters: host copies, then one upload (the preprocess takes its
    // views' parameters and the scene as kernel arguments; the other kernels
    // read them from this block, so a captured graph replays with new ones)
    const size_t vp_stride = align_up(sizeof(ViewParams));
    const size_t pbytes = vp_stride * (size_t)n_views;
    int rc = ensure(h, h->params, pbytes);
    if (rc) return rc;
    if (h->params_host_cap < pbytes) {
        std::lock_guard<std::mutex> lock(device_memory_lock());
        if (h->params_host) HCHK(hipHostFree(h->params_host));
        h->params_host = nullptr;
        h->params_host_cap = 0;
        HCHK(hipHostMalloc(&h->params_host, h->params.cap, hipHostMallocDefault));
        h->params_host_cap = h->params.cap;
    }
    char* ph = (char*)h->params_host;
    char* pd = (char*)h->params.ptr;
    for (int v = 0; v < n_views; ++v) {
        ViewSlot& slot = (*h->cur[v]);
        const fhtgs_scene& sc = scenes[view_scene[v]];
        fill_params(sc, views[v], slot.vp);
        slot.dvp = (ViewParams*)(pd + vp_stride * (size_t)v);
        const int64_t N = sc.n_primitives;
        // a view without summed-area tables gets them built into its slot
        slot.auto_tables = views[v].fovea_mask_area_table == nullptr;
        const int tR = slot.auto_tables ? slot.vp.fovea_r : -1;
        size_t bytes = 0;
        carve_fixed(nullptr, N, slot.vp.n_large, slot.vp.n_small, slot.vb, bytes, slot.vp.gw, slot.vp.gh, tR);
        if ((rc = ensure(h, slot.fixed, bytes))) return rc;
        carve_fixed(slot.fixed.ptr, N, slot.vp.n_large, slot.vp.n_small, slot.vb, bytes, slot.vp.gw, slot.vp.gh, tR);
        if (slot.auto_tables) {
            if (slot.vp.n_large > kMaxLargeTiles) return set_err(h, FHTGS_E_UNSUPPORTED, "image too large for device-built tables");
            slot.own_mask = views[v].visibility_mask == nullptr;
            if (slot.own_mask) slot.vp.vis_mask = slot.vb.tab_mask;
            slot.vp.vis_sat = slot.vb.tab_vis_sat;
            slot.vp.fovea_sat = slot.vb.tab_fovea_sat;
        }
        std::memcpy(ph + vp_stride * (size_t)v, &slot.vp, sizeof(ViewParams));
        if (!slot.host_fi) {
            std::lock_guard<std::mutex> lock(device_memory_lock());
            HCHK(hipHostMalloc((void**)&slot.host_fi, FI_WORDS * sizeof(uint32_t), hipHostMallocDefault));
        }
        slot.n_prims = N;
    }

    // ---- mode: device counts need a capacity from an exact frame of this shape
    int mode = force_mode >= 0 ? force_mode : h->frame_mode;
    if (h->keep_unsorted) mode = kModeExact;
    for (int v = 0; v < n_views && mode != kModeExact; ++v) {
        const ViewSlot& slot = (*h->cur[v]);
        if (slot.cap_hint <= 0 || slot.cap_for_n != slot.n_prims || slot.cap_for_tiles != slot.vp.n_small)
            mode = kModeExact;
    }
    if (mode != kModeExact) {
        for (int v = 0; v < n_views; ++v)
            if ((rc = size_instances(h, (*h->cur[v]), (*h->cur[v]).cap_hint))) return rc;
    }
    const bool timed = h->timing != 0 && mode != kModeGraph;

    if (mode == kModeExact) {
        if ((rc = enqueue_phase1(h, groups, scenes, pbytes, n_views, timed, false, s))) return rc;
        for (int v = 0; v < n_views; ++v)
            HCHK(hipMemcpyAsync((*h->cur[v]).host_fi, (*h->cur[v]).vb.frame_info, FI_WORDS * sizeof(uint32_t),
                                hipMemcpyDeviceToHost, s));
        if (timed && (rc = record(h, 4, s))) return rc;
        HCHK(hipStreamSynchronize(s));
        for (int v = 0; v < n_views; ++v) {
            ViewSlot& slot = (*h->cur[v]);
            if (slot.host_fi[FI_ERROR] & 8u) return set_err(h, FHTGS_E_TABLES, kTablesMismatch);
            const uint32_t I = slot.n_prims > 0 ? slot.host_fi[FI_INSTANCES] : 0u;
            if (I >= (1u << 30)) return set_err(h, FHTGS_E_INVALID, "more than 2^30 tile instances in one view");
            slot.n_instances = I;
            slot.n_active = slot.host_fi[FI_N_ACTIVE];
            if ((rc = size_instances(h, slot, I))) return rc;
            if (h->keep_unsorted) {
                const size_t db = align_up((size_t)std::max<uint32_t>(I, 1) * 4) * 2;
                if ((rc = ensure(h, slot.debug, db))) return rc;
                slot.dbg_keys = (uint32_t*)slot.debug.ptr;
                slot.dbg_vals = (uint32_t*)((char*)slot.debug.ptr + db / 2);
            }
        }
        if ((rc = enqueue_phase2(h, views, n_views, true, timed, s))) return rc;
    } else if (mode == kModeDevice) {
        if ((rc = enqueue_phase1(h, groups, scenes, pbytes, n_views, timed, true, s))) return rc;
        if (timed && (rc = record(h, 4, s))) return rc;
        if ((rc = enqueue_phase2(h, views, n_views, false, timed, s))) return rc;
    } else {
        if (h->timing) HCHK(hipEventRecord(h->ev[0], s));
        // phase 1 straight onto the stream (its first kernels start at once);
        // the captured graph holds phase 2 and is launched while phase 1 runs
        if ((rc = enqueue_phase1(h, groups, scenes, pbytes, n_views, false, true, s))) return rc;
        std::vector<int64_t> sig = launch_signature(h, groups, scenes, views, n_views);
        GraphEntry* ge = nullptr;
        for (GraphEntry& e : h->graphs)
            if (e.sig == sig) ge = &e;
        if (!ge) {
            std::lock_guard<std::mutex> lock(device_memory_lock());
            if (h->graphs.size() >= kMaxGraphs) {
                auto lru = std::min_element(h->graphs.begin(), h->graphs.end(),
                                            [](const GraphEntry& x, const GraphEntry& y) { return x.last_use < y.last_use; });
                HCHK(hipGraphExecDestroy(lru->exec));
                h->graphs.erase(lru);
            }
            HCHK(hipStreamBeginCapture(h->cap_stream, hipStreamCaptureModeThreadLocal));
            int crc = enqueue_phase2(h, views, n_views, false, false, h->cap_stream);
            hipGraph_t graph = nullptr;
            const hipError_t ce = hipStreamEndCapture(h->cap_stream, &graph);
            if (crc) {
                if (graph) (void)hipGraphDestroy(graph);
                return crc;
            }
            if (ce != hipSuccess) return set_err(h, FHTGS_E_HIP, std::string("graph capture: ") + hipGetErrorString(ce));
            GraphEntry e;
            e.sig = sig;
            const hipError_t ie = hipGraphInstantiate(&e.exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (ie != hipSuccess) return set_err(h, FHTGS_E_HIP, std::string("graph instantiate: ") + hipGetErrorString(ie));
            h->graphs.push_back(e);
            ge = &h->graphs.back();
        }
        ge->last_use = ++h->graph_clock;
        HCHK(hipGraphLaunch(ge->exec, s));
        for (int v = 0; v < n_views; ++v) (*h->cur[v]).sorted_src = (*h->cur[v]).plan.passes & 1;
        if (h->timing) HCHK(hipEventRecord(h->ev[FHTGS_NUM_STAGES - 1], s));
    }
    // frame info at the end: counts, valid and visible totals, error bits
    for (int v = 0; v < n_views; ++v)
        HCHK(hipMemcpyAsync((*h->cur[v]).host_fi, (*h->cur[v]).vb.frame_info, FI_WORDS * sizeof(uint32_t),
                            hipMemcpyDeviceToHost, s));
    HCHK(hipEventRecord(h->done, s));
    h->have_done = true;
    // the reference call returns after its streams finish (inference.cu:468-470)
    HCHK(hipStreamSynchronize(s));
    h->last_mode = mode;

    bool redo = false;
    for (int v = 0; v < n_views; ++v) {
        ViewSlot& slot = (*h->cur[v]);
        const uint32_t* fi = slot.host_fi;
        if (fi[FI_ERROR] & 8u) return set_err(h, FHTGS_E_TABLES, kTablesMismatch);
        if (fi[FI_ERROR] & 1u)
            return set_err(h, FHTGS_E_TABLES,
                           "tiles-touched from the culling tables under-counts the visible tiles of a Gaussian");
        if (fi[FI_ERROR] & 4u) {
            redo = true;
            continue;
        }
        if (fi[FI_ERROR] & (uint32_t)kSortErrorBit)
            return set_err(h, FHTGS_E_HIP, "instance sort detected an inconsistency (look-back or scatter)");
        const uint32_t I = slot.n_prims > 0 ? fi[FI_INSTANCES] : 0u;
        slot.n_instances = I;
        slot.n_active = fi[FI_N_ACTIVE];
        if (mode == kModeExact || (int64_t)I > slot.ib.capacity - slot.ib.capacity / 8) {
            slot.cap_hint = with_headroom(I);
            slot.cap_for_n = slot.n_prims;
            slot.cap_for_tiles = slot.vp.n_small;
        }
        int64_t* c = slot.counts;
        c[0] = slot.n_prims; c[1] = slot.vp.n_large; c[2] = slot.vp.n_small; c[3] = slot.n_active;
        c[4] = I; c[5] = fi[FI_VALID]; c[6] = slot.plan.end_bit; c[7] = slot.plan.key_bytes;
        c[8] = fi[FI_N_VISIBLE];
        c[9] = slot.vp.gw; c[10] = slot.vp.gh; c[11] = slot.vp.fovea_r; c[12] = slot.vp.blend_r;
        c[13] = slot.vp.gtx; c[14] = slot.vp.gty; c[15] = slot.plan.passes;
    }
    if (redo) {
        // a view outgrew its instance capacity: render the frame again with
        // exact buffers (the next frames use the new capacity)
        return render_impl(h, scenes, n_scenes, views, view_scene, view_slot, n_views, pair_views, s, kModeExact);
    }
    if (h->timing) {
        // graph frames time the whole frame only
        for (int i = 0; i < FHTGS_NUM_STAGES - 1; ++i) {
            float ms = 0.0f;
            if (mode != kModeGraph) HCHK(hipEventElapsedTime(&ms, h->ev[i], h->ev[i + 1]));
            h->stage_ms[i] = ms;
        }
        float tot = 0.0f;
        HCHK(hipEventElapsedTime(&tot, h->ev[0], h->ev[FHTGS_NUM_STAGES - 1]));
        h->stage_ms[FHTGS_NUM_STAGES - 1] = tot;
    }
    return FHTGS_OK;
}

}  // namespace

extern "C" {

const char* fhtgs_version(void) { return "fhtgs-gfx950 0.1"; }

int fhtgs_create(int device, fhtgs_handle* out) {
    if (!out) return FHTGS_E_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) {
        (void)hipGetLastError();
        return FHTGS_E_INVALID;
    }
    fhtgs_context* h = new fhtgs_context();
    h->device = device;
    int prev = 0;
    (void)hipGetDevice(&prev);
    if (hipSetDevice(device) != hipSuccess || hipEventCreateWithFlags(&h->done, hipEventDisableTiming) != hipSuccess) {
        delete h;
        return FHTGS_E_HIP;
    }
    for (int i = 0; i <= FHTGS_NUM_STAGES; ++i) {
        if (hipEventCreate(&h->ev[i]) != hipSuccess) {
            delete h;
            return FHTGS_E_HIP;
        }
    }
    bool ok = hipStreamCreateWithFlags(&h->cap_stream, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&h->ev_params, hipEventDisableTiming) == hipSuccess;
    for (int b = 0; ok && b < kMaxBranches - 1; ++b)
        ok = hipStreamCreateWithFlags(&h->side[b], hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&h->ev_join[b], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        (void)fhtgs_destroy(h);
        (void)hipSetDevice(prev);
        return FHTGS_E_HIP;
    }
    (void)hipSetDevice(prev);
    *out = h;
    return FHTGS_OK;
}

int fhtgs_destroy(fhtgs_handle h) {
    if (!h) return FHTGS_OK;
    int prev = 0;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(h->device);
    if (h->have_done) (void)hipEventSynchronize(h->done);
    std::lock_guard<std::mutex> lock(device_memory_lock());
    for (auto& s : h->slots) {
        if (s.fixed.ptr) (void)hipFree(s.fixed.ptr);
        if (s.inst.ptr) (void)hipFree(s.inst.ptr);
        if (s.debug.ptr) (void)hipFree(s.debug.ptr);
        if (s.host_fi) (void)hipHostFree(s.host_fi);
    }
    if (h->sort_scratch.ptr) (void)hipFree(h->sort_scratch.ptr);
    if (h->vis.fixed.ptr) (void)hipFree(h->vis.fixed.ptr);
    if (h->vis_params.ptr) (void)hipFree(h->vis_params.ptr);
    if (h->vis_params_host) (void)hipHostFree(h->vis_params_host);
    if (h->params.ptr) (void)hipFree(h->params.ptr);
    if (h->params_host) (void)hipHostFree(h->params_host);
    for (auto& g : h->graphs) (void)hipGraphExecDestroy(g.exec);
    if (h->cap_stream) (void)hipStreamDestroy(h->cap_stream);
    for (int b = 0; b < kMaxBranches - 1; ++b) {
        if (h->side[b]) (void)hipStreamDestroy(h->side[b]);
        if (h->ev_join[b]) (void)hipEventDestroy(h->ev_join[b]);
    }
    if (h->ev_fork) (void)hipEventDestroy(h->ev_fork);
    if (h->ev_params) (void)hipEventDestroy(h->ev_params);
    if (h->done) (void)hipEventDestroy(h->done);
    for (int i = 0; i <= FHTGS_NUM_STAGES; ++i)
        if (h->ev[i]) (void)hipEventDestroy(h->ev[i]);
    (void)hipSetDevice(prev);
    delete h;
    return FHTGS_OK;
}

const char* fhtgs_last_error(fhtgs_handle h) {
    if (h) return h->err.c_str();
    return g_no_handle_err.empty() ? "null handle" : g_no_handle_err.c_str();
}

int fhtgs_fovea_radius_tiles(int32_t width, int32_t height, int32_t fr_px, int32_t bw_px,
                             int32_t* fovea_tiles, int32_t* blend_tiles) {
    if (width <= 0 || height <= 0 || !fovea_tiles || !blend_tiles) return FHTGS_E_INVALID;
    int R, B;
    fov_radii(width, height, fr_px, bw_px, R, B);
    *fovea_tiles = R;
    *blend_tiles = B;
    return FHTGS_OK;
}

int fhtgs_set_option(fhtgs_handle h, int32_t option, int64_t value) {
    if (!h) return FHTGS_E_INVALID;
    switch (option) {
        case FHTGS_OPT_KEEP_UNSORTED: h->keep_unsorted = value ? 1 : 0; return FHTGS_OK;
        case FHTGS_OPT_SORT_DIGIT_BITS:
            if (value < 4 || value > 11) return set_err(h, FHTGS_E_INVALID, "digit bits must be 4..11");
            h->digit_bits = (int)value;
            return FHTGS_OK;
        case FHTGS_OPT_STAGE_TIMING: h->timing = value ? 1 : 0; return FHTGS_OK;
        case FHTGS_OPT_FRAME_MODE:
            if (value < 0 || value > 2) return set_err(h, FHTGS_E_INVALID, "frame mode must be 0, 1 or 2");
            h->frame_mode = (int)value;
            return FHTGS_OK;
        case FHTGS_OPT_VIEW_BRANCHES:
            if (value < 1 || value > kMaxBranches) return set_err(h, FHTGS_E_INVALID, "view branches must be 1..4");
            h->branches = (int)value;
            return FHTGS_OK;
        case FHTGS_OPT_BLEND_XCD_RUN:
            if (value < 0 || value > 4096) return set_err(h, FHTGS_E_INVALID, "XCD run must be 0..4096");
            h->xcd_run = (int)value;
            return FHTGS_OK;
        case FHTGS_OPT_BLEND_WORK_TILE:
            if (value != 1 && value != 2 && value != 4)
                return set_err(h, FHTGS_E_INVALID, "blend work tile must be 1 (8x8), 2 (16x8) or 4 (16x16)");
            h->work_tile = (int)value;
            h->tuned_wt.clear();
            return FHTGS_OK;
        default: return set_err(h, FHTGS_E_INVALID, "unknown option");
    }
}

int fhtgs_workspace_bytes(fhtgs_handle h, int64_t* bytes) {
    if (!h || !bytes) return FHTGS_E_INVALID;
    int64_t t = (int64_t)(h->sort_scratch.cap + h->params.cap);
    for (auto& s : h->slots) t += (int64_t)(s.fixed.cap + s.inst.cap + s.debug.cap);
    *bytes = t;
    return FHTGS_OK;
}

int fhtgs_workspace_query(const fhtgs_view* view, int64_t n_primitives, int64_t n_instances,
                          int64_t* fixed_bytes, int64_t* instance_bytes) {
    if (!view || !fixed_bytes || !instance_bytes || n_primitives < 0 || n_instances < 0)
        return set_err(nullptr, FHTGS_E_INVALID, "null view or output, or a negative count");
    fhtgs_context probe;  // host-side only: collects the message, touches no device
    if (const int rc = validate_geometry(&probe, *view, "view: ")) return set_err(nullptr, rc, probe.err);
    fhtgs_scene sc{};
    sc.n_primitives = n_primitives;
    ViewParams vp{};
    fill_params(sc, *view, vp);
    // the same carving a render does: per-Gaussian and per-tile buffers (with
    // device-built culling tables when the view brings none), then the
    // instance buffers under the largest sort plan any digit width gives
    const int tR = view->fovea_mask_area_table == nullptr ? vp.fovea_r : -1;
    ViewBuffers vb{};
    size_t fixed = 0;
    carve_fixed(nullptr, n_primitives, vp.n_large, vp.n_small, vb, fixed, vp.gw, vp.gh, tR);
    const int end_bit = end_bit_of((uint32_t)vp.n_small + 1u);
    size_t inst = 0;
    for (int bits = 4; bits <= 11; ++bits) {
        InstanceBuffers ib{};
        size_t b = 0;
        carve_inst(nullptr, n_instances, make_plan(end_bit, bits), ib, b);
        inst = std::max(inst, b);
    }
    *fixed_bytes = (int64_t)fixed;
    *instance_bytes = (int64_t)inst;
    return FHTGS_OK;
}

int fhtgs_stage_times(fhtgs_handle h, float* ms) {
    if (!h || !ms) return FHTGS_E_INVALID;
    for (int i = 0; i < FHTGS_NUM_STAGES; ++i) ms[i] = h->stage_ms[i];
    return FHTGS_OK;
}

namespace {

int check_scene(fhtgs_context* h, const fhtgs_scene& sc, int idx) {
    const std::string p = "scene " + std::to_string(idx) + ": ";
    if (sc.n_primitives < 0 || sc.n_primitives >= ((int64_t)1 << 31))
        return set_err(h, FHTGS_E_INVALID, p + "primitive count out of range");
    if (sc.n_primitives > 0 && (!sc.positions || !sc.scales || !sc.rotations || !sc.opacities || !sc.sh_0 ||
                                (!sc.sh_rest && sc.total_sh_bases > 0)))
        return set_err(h, FHTGS_E_INVALID, p + "null scene tensor");
    return FHTGS_OK;
}

int validate_call(fhtgs_context* h, const fhtgs_scene* scenes, int n_scenes, const fhtgs_view* views,
                  const int32_t* view_scene, const int32_t* view_slot, int n_views) {
    for (int k = 0; k < n_scenes; ++k) {
        const int rc = check_scene(h, scenes[k], k);
        if (rc) return rc;
    }
    std::vector<char> used;
    for (int v = 0; v < n_views; ++v) {
        const int rc = validate_view(h, views[v], v);
        if (rc) return rc;
        if (view_scene[v] < 0 || view_scene[v] >= n_scenes)
            return set_err(h, FHTGS_E_INVALID, "view " + std::to_string(v) + ": scene index out of range");
        // the degree that active_sh_bases selects reads 3, 8 or 15 coefficients
        // per Gaussian (sh_utils.cuh:32-69); fewer given would be read past
        const int act = views[v].active_sh_bases;
        const int need = act <= 1 ? 0 : (act <= 4 ? 3 : (act <= 9 ? 8 : 15));
        if (need > scenes[view_scene[v]].total_sh_bases)
            return set_err(h, FHTGS_E_INVALID, "active_sh_bases exceeds the SH coefficients given");
        if (view_slot) {
            const int sl = view_slot[v];
            if (sl < 0 || sl >= kMaxSlots)
                return set_err(h, FHTGS_E_INVALID, "view " + std::to_string(v) + ": slot out of range");
            if ((int)used.size() <= sl) used.resize((size_t)sl + 1, 0);
            if (used[(size_t)sl]) return set_err(h, FHTGS_E_INVALID, "two views of one call share a slot");
            used[(size_t)sl] = 1;
        }
    }
    return FHTGS_OK;
}

struct DeviceScope {
    int prev = 0, dev = 0;
    bool ok = false;
    DeviceScope(int d) : dev(d) {
        ok = hipGetDevice(&prev) == hipSuccess && (prev == d || hipSetDevice(d) == hipSuccess);
    }
    ~DeviceScope() {
        if (ok && prev != dev) (void)hipSetDevice(prev);
    }
};

int render_entry(fhtgs_context* h, const fhtgs_scene* scenes, int n_scenes, const fhtgs_view* views,
                 const int32_t* view_scene, const int32_t* view_slot, int n_views, bool pair_views,
                 hipStream_t s, int force_mode) {
    int rc = validate_call(h, scenes, n_scenes, views, view_scene, view_slot, n_views);
    if (rc) return rc;
    DeviceScope scope(h->device);
    if (!scope.ok) return set_err(h, FHTGS_E_HIP, "cannot select the handle's device");
    if (h->have_done) HCHK(hipStreamWaitEvent(s, h->done, 0));
    return render_impl(h, scenes, n_scenes, views, view_scene, view_slot, n_views, pair_views, s, force_mode);
}

}  // namespace

int fhtgs_validate_view(const fhtgs_view* view, char* msg, int32_t msg_len) {
    if (msg && msg_len > 0) msg[0] = '\0';
    if (!view) return FHTGS_E_INVALID;
    fhtgs_context probe;  // host-side only: collects the message, touches no device
    const int rc = validate_view(&probe, *view, 0);
    if (rc && msg && msg_len > 0) std::snprintf(msg, (size_t)msg_len, "%s", probe.err.c_str());
    return rc;
}

int fhtgs_render(fhtgs_handle h, const fhtgs_scene* scene, const fhtgs_view* views, int32_t n_views,
                 int32_t share_preprocess, void* stream_ptr) {
    if (!h) return FHTGS_E_INVALID;
    if (!scene || !views || n_views <= 0) return set_err(h, FHTGS_E_INVALID, "null scene/views or no view");
    if (share_preprocess && n_views > 2)
        return set_err(h, FHTGS_E_INVALID, "a shared preprocess covers at most two views");
    if (share_preprocess && n_views == 2 && !same_preprocess(views[0], views[1]))
        return set_err(h, FHTGS_E_INVALID, "shared-preprocess views disagree on preprocess parameters");
    std::vector<int32_t> vs((size_t)n_views, 0);
    return render_entry(h, scene, 1, views, vs.data(), nullptr, n_views, share_preprocess != 0,
                        (hipStream_t)stream_ptr, -1);
}

int fhtgs_render_batch(fhtgs_handle h, const fhtgs_scene* scenes, int32_t n_scenes, const fhtgs_view* views,
                       const int32_t* view_scene, const int32_t* view_slot, int32_t n_views, void* stream_ptr) {
    if (!h) return FHTGS_E_INVALID;
    if (!scenes || n_scenes <= 0 || !views || !view_scene || n_views <= 0)
        return set_err(h, FHTGS_E_INVALID, "null scenes/views/view_scene or empty batch");
    return render_entry(h, scenes, n_scenes, views, view_scene, view_slot, n_views, true,
                        (hipStream_t)stream_ptr, -1);
}

int fhtgs_autotune(fhtgs_handle h, const fhtgs_scene* scenes, int32_t n_scenes, const fhtgs_view* views,
                   const int32_t* view_scene, const int32_t* view_slot, int32_t n_views, int32_t frames,
                   void* stream_ptr) {
    if (!h) return FHTGS_E_INVALID;
    if (!scenes || n_scenes <= 0 || !views || !view_scene || n_views <= 0 || frames <= 0)
        return set_err(h, FHTGS_E_INVALID, "null scenes/views/view_scene, empty batch or no frames");
    hipStream_t s = (hipStream_t)stream_ptr;
    // exact frame first: sizes every slot's buffers and instance capacity
    int rc = render_entry(h, scenes, n_scenes, views, view_scene, view_slot, n_views, true, s, kModeExact);
    if (rc) return rc;
    DeviceScope scope(h->device);
    if (!scope.ok) return set_err(h, FHTGS_E_HIP, "cannot select the handle's device");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HCHK(hipEventCreate(&e0));
    HCHK(hipEventCreate(&e1));
    auto timed = [&](const std::vector<int32_t>& sel, float& ms) -> int {
        // device-count frames of the selected views, best of `frames`
        std::vector<fhtgs_view> vv;
        std::vector<int32_t> vsc, vsl;
        for (int v : sel) {
            vv.push_back(views[v]);
            vsc.push_back(view_scene[v]);
            vsl.push_back(view_slot ? view_slot[v] : v);
        }
        ms = 1e30f;
        for (int f = 0; f <= frames; ++f) {
            HCHK(hipEventRecord(e0, s));
            int r = render_impl(h, scenes, n_scenes, vv.data(), vsc.data(), vsl.data(), (int)vv.size(), true, s,
                                kModeDevice);
            if (r) return r;
            HCHK(hipEventRecord(e1, s));
            HCHK(hipEventSynchronize(e1));
            float t = 0.0f;
            HCHK(hipEventElapsedTime(&t, e0, e1));
            if (f > 0) ms = std::min(ms, t);  // the first run is a warm-up
        }
        return FHTGS_OK;
    };
    // sort digit width per view size: 8-bit digits (more, cheaper passes) or
    // 11-bit digits (fewest passes); images do not depend on it
    std::vector<int64_t> sizes;
    for (int v = 0; v < n_views; ++v) {
        const int64_t ns = (int64_t)((views[v].width + kTileL - 1) / kTileL) * ((views[v].height + kTileL - 1) / kTileL) * kNumSub;
        if (std::find(sizes.begin(), sizes.end(), ns) == sizes.end()) sizes.push_back(ns);
    }
    for (int64_t ns : sizes) {
        std::vector<int32_t> sel;
        for (int v = 0; v < n_views; ++v) {
            const int64_t vs = (int64_t)((views[v].width + kTileL - 1) / kTileL) * ((views[v].height + kTileL - 1) / kTileL) * kNumSub;
            if (vs == ns) sel.push_back(v);
        }
        int best_bits = 11;
        float best = 1e30f;
        for (int bits : {8, 11}) {
            auto it = std::find_if(h->tuned_bits.begin(), h->tuned_bits.end(), [&](const std::pair<int64_t, int>& t) { return t.first == ns; });
            if (it == h->tuned_bits.end()) h->tuned_bits.push_back({ns, bits});
            else it->second = bits;
            float ms = 0.0f;
            if ((rc = timed(sel, ms))) break;
            if (ms < best) { best = ms; best_bits = bits; }
        }
        if (rc) break;
        for (auto& t : h->tuned_bits)
            if (t.first == ns) t.second = best_bits;
        // the blend's work tile for this size: 8x8, 16x8 or 16x16 (waves per
        // workgroup 1, 2, 4); every pixel sees the same Gaussians in the same
        // order with each
        int best_wt = work_tile_for(h, ns);
        best = 1e30f;
        for (int wt : {1, 2, 4}) {
            auto it = std::find_if(h->tuned_wt.begin(), h->tuned_wt.end(), [&](const std::pair<int64_t, int>& t) { return t.first == ns; });
            if (it == h->tuned_wt.end()) h->tuned_wt.push_back({ns, wt});
            else it->second = wt;
            float ms = 0.0f;
            if ((rc = timed(sel, ms))) break;
            if (ms < best) { best = ms; best_wt = wt; }
        }
        for (auto& t : h->tuned_wt)
            if (t.first == ns) t.second = best_wt;
        if (rc) break;
    }
    std::vector<int32_t> all;
    for (int v = 0; v < n_views; ++v) all.push_back(v);
    // blend tile runs per XCD, over the whole batch
    if (!rc) {
        int best_run = h->xcd_run;
        float best = 1e30f;
        for (int run : {0, 16, 64}) {
            h->xcd_run = run;
            float ms = 0.0f;
            if ((rc = timed(all, ms))) break;
            if (ms < best) { best = ms; best_run = run; }
        }
        h->xcd_run = best_run;
    }
    // concurrent view branches, over the whole batch: more branches overlap
    // more of the views' latency-bound instance and sort kernels, and compete
    // for the same CUs in their blends
    if (!rc && n_views > 1) {
        const int keep = h->branches;
        int best_nb = keep;
        float best = 1e30f;
        for (int nb = 2; nb <= kMaxBranches; ++nb) {
            h->branches = nb;
            float ms = 0.0f;
            if ((rc = timed(all, ms))) break;
            if (ms < best) { best = ms; best_nb = nb; }
        }
        h->branches = rc ? keep : best_nb;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

int fhtgs_get_option(fhtgs_handle h, int32_t option, int64_t* value) {
    if (!h || !value) return FHTGS_E_INVALID;
    switch (option) {
        case FHTGS_OPT_KEEP_UNSORTED: *value = h->keep_unsorted; return FHTGS_OK;
        case FHTGS_OPT_SORT_DIGIT_BITS: *value = h->digit_bits; return FHTGS_OK;
        case FHTGS_OPT_STAGE_TIMING: *value = h->timing; return FHTGS_OK;
        case FHTGS_OPT_BLEND_XCD_RUN: *value = h->xcd_run; return FHTGS_OK;
        case FHTGS_OPT_FRAME_MODE: *value = h->frame_mode; return FHTGS_OK;
        case FHTGS_OPT_VIEW_BRANCHES: *value = h->branches; return FHTGS_OK;
        case FHTGS_OPT_BLEND_WORK_TILE: *value = h->work_tile; return FHTGS_OK;
        default: return set_err(h, FHTGS_E_INVALID, "unknown option");
    }
}

int fhtgs_tuned_digit_bits(fhtgs_handle h, int32_t width, int32_t height, int32_t* bits) {
    if (!h || !bits || width <= 0 || height <= 0) return FHTGS_E_INVALID;
    const int64_t ns = (int64_t)((width + kTileL - 1) / kTileL) * ((height + kTileL - 1) / kTileL) * kNumSub;
    *bits = digit_bits_for(h, ns);
    return FHTGS_OK;
}

int fhtgs_tuned_work_tile(fhtgs_handle h, int32_t width, int32_t height, int32_t* waves) {
    if (!h || !waves || width <= 0 || height <= 0) return FHTGS_E_INVALID;
    const int64_t ns = (int64_t)((width + kTileL - 1) / kTileL) * ((height + kTileL - 1) / kTileL) * kNumSub;
    *waves = work_tile_for(h, ns);
    return FHTGS_OK;
}

int fhtgs_build_tables(fhtgs_handle h, const uint8_t* visible, int32_t width, int32_t height,
                       int32_t foveation_radius_px, int32_t blend_width_px, uint32_t* mask_out,
                       uint32_t* vis_sat_out, uint32_t* fovea_sat_out, void* stream_ptr) {
    if (!h) return FHTGS_E_INVALID;
    if (width <= 0 || height <= 0) return set_err(h, FHTGS_E_INVALID, "width and height must be positive");
    if (!mask_out && vis_sat_out) return set_err(h, FHTGS_E_INVALID, "the visibility SAT needs the mask output");
    const int gw = (width + kTileL - 1) / kTileL, gh = (height + kTileL - 1) / kTileL;
    if (gw * gh > kMaxLargeTiles) return set_err(h, FHTGS_E_UNSUPPORTED, "more than 65536 tiles");
    int R = 0, B = 0;
    fov_radii(width, height, foveation_radius_px, blend_width_px, R, B);
    DeviceScope scope(h->device);
    if (!scope.ok) return set_err(h, FHTGS_E_HIP, "cannot select the handle's device");
    HCHK(launch_build_tables(visible, mask_out != nullptr, gw, gh, R, mask_out, vis_sat_out, fovea_sat_out,
                             (hipStream_t)stream_ptr));
    return FHTGS_OK;
}

int fhtgs_visualize_gaze(fhtgs_handle h, float* image, int32_t width, int32_t height, float gaze_x,
                         float gaze_y, const float* rgb, int32_t type, int32_t size, int32_t to_chw,
                         void* stream_ptr) {
    if (!h) return FHTGS_E_INVALID;
    if (!image || !rgb || width <= 0 || height <= 0) return set_err(h, FHTGS_E_INVALID, "null image/colour or bad size");
    if (type < 0 || type > 2) return set_err(h, FHTGS_E_INVALID, "visualization type must be 0, 1 or 2");
    DeviceScope scope(h->device);
    if (!scope.ok) return set_err(h, FHTGS_E_HIP, "cannot select the handle's device");
    hipStream_t s = (hipStream_t)stream_ptr;
    if (h->have_done) HCHK(hipStreamWaitEvent(s, h->done, 0));
    HCHK(launch_gaze_marker(image, width, height, gaze_x, gaze_y, rgb, size, type, to_chw ? 1 : 0, s));
    HCHK(hipEventRecord(h->done, s));
    h->have_done = true;
    return FHTGS_OK;
}

int fhtgs_visualize_tile_boundaries(fhtgs_handle h, float* image, int32_t width, int32_t height,
                                    const uint32_t* visibility_mask, float gaze_x, float gaze_y,
                                    int32_t foveation_radius_px, int32_t blend_width_px, int32_t to_chw,
                                    void* stream_ptr) {
    if (!h) return FHTGS_E_INVALID;
    if (!image || !visibility_mask || width <= 0 || height <= 0)
        return set_err(h, FHTGS_E_INVALID, "null image/mask or bad size");
    if (foveation_radius_px >= 0 && blend_width_px > foveation_radius_px)
        return set_err(h, FHTGS_E_INVALID, "blend width exceeds the foveation radius");
    if (foveation_radius_px >= 0 && blend_width_px <= kTileL)
        return set_err(h, FHTGS_E_INVALID, "blend width must exceed 16 px with foveation on");
    DeviceScope scope(h->device);
    if (!scope.ok) return set_err(h, FHTGS_E_HIP, "cannot select the handle's device");
    hipStream_t s = (hipStream_t)stream_ptr;
    if (h->have_done) HCHK(hipEventSynchronize(h->done));  // staging and tile buffers are reused
    fhtgs_view v;
    std::memset(&v, 0, sizeof(v));
    v.width = width;
    v.height = height;
    v.gaze_x = gaze_x;
    v.gaze_y = gaze_y;
    v.visibility_mask = visibility_mask;
    v.foveation_radius_px = foveation_radius_px;
    v.blend_width_px = blend_width_px;
    v.image = image;
    fhtgs_scene sc;
    std::memset(&sc, 0, sizeof(sc));
    ViewSlot& sl = h->vis;
    fill_params(sc, v, sl.vp);
    if (sl.vp.n_large > kMaxLargeTiles) return set_err(h, FHTGS_E_UNSUPPORTED, "image too large for the tile overlay");
    size_t bytes = 0;
    carve_fixed(nullptr, 0, sl.vp.n_large, sl.vp.n_small, sl.vb, bytes);
    int rc = ensure(h, sl.fixed, bytes);
    if (rc) return rc;
    carve_fixed(sl.fixed.ptr, 0, sl.vp.n_large, sl.vp.n_small, sl.vb, bytes);
    if ((rc = ensure(h, h->vis_params, sizeof(ViewParams)))) return rc;
    if (!h->vis_params_host) {
        std::lock_guard<std::mutex> lock(device_memory_lock());
        HCHK(hipHostMalloc((void**)&h->vis_params_host, sizeof(ViewParams), hipHostMallocDefault));
    }
    *h->vis_params_host = sl.vp;
    const ViewParams* dvp = (const ViewParams*)h->vis_params.ptr;
    if (h->have_done) HCHK(hipStreamWaitEvent(s, h->done, 0));
    HCHK(hipMemcpyAsync(h->vis_params.ptr, h->vis_params_host, sizeof(ViewParams), hipMemcpyHostToDevice, s));
    // the rasterizer's own tile classification (visualization.cu:60-178)
    HCHK(launch_clear_frame(sl.vp, sl.vb, s));
    HCHK(launch_tiles(sl.vp, dvp, sl.vb, s));
    HCHK(launch_tile_borders(sl.vp, dvp, sl.vb, image, to_chw ? 1 : 0, s));
    HCHK(hipEventRecord(h->done, s));
    h->have_done = true;
    // the reference call returns after its streams finish (visualization.cu:232-234)
    HCHK(hipStreamSynchronize(s));
    return FHTGS_OK;
}

int fhtgs_debug_counts(fhtgs_handle h, int32_t view, int64_t* counts) {
    if (!h || !counts || view < 0 || view >= (int)h->slots.size()) return FHTGS_E_INVALID;
    std::memcpy(counts, h->slots[view].counts, sizeof(int64_t) * 16);
    return FHTGS_OK;
}

int fhtgs_debug_read(fhtgs_handle h, int32_t view, int32_t id, void* dst, int64_t dst_bytes,
                     int64_t* written) {
    if (!h || view < 0 || view >= (int)h->slots.size() || !dst) return FHTGS_E_INVALID;
    ViewSlot& sl = h->slots[view];
    const int64_t N = sl.n_prims, I = sl.n_instances, na = sl.n_active;
    const void* src = nullptr;
    int64_t bytes = 0;
    bool widen = false;
    int field = -1;  // >= 0: this float4 of each 64-B primitive record
    switch (id) {
        case FHTGS_DBG_TILE_NUM: src = sl.vb.tile_num; bytes = (int64_t)sl.vp.n_large * 4; break;
        case FHTGS_DBG_TILE_MAP: src = sl.vb.tile_map; bytes = na * 4; break;
        case FHTGS_DBG_TILE_TYPE: src = sl.vb.tile_type; bytes = na; break;
        case FHTGS_DBG_TILE_MAP_PART: src = sl.vb.tile_map_part; bytes = na * 4; break;
        case FHTGS_DBG_PARTITION: src = sl.vb.frame_info + FI_PART; bytes = 24; break;
        case FHTGS_DBG_PARTITION_QUARTER: src = sl.vb.frame_info + FI_PART + 6; bytes = 8; break;
        case FHTGS_DBG_BLEND_ORDER:
            src = sl.vb.blend_order;
            bytes = (int64_t)blend_units(sl.host_fi + FI_PART, (uint32_t)na, work_tile_for(h, sl.vp.n_small)) * 4;
            break;
        case FHTGS_DBG_N_TOUCHED: src = sl.vb.n_touched; bytes = N * 4; break;
        case FHTGS_DBG_BOUNDS: src = sl.vb.bounds; bytes = N * 8; break;
        case FHTGS_DBG_VPMT1: src = sl.vb.prim; bytes = N * 16; field = 0; break;
        case FHTGS_DBG_VPMT2: src = sl.vb.prim; bytes = N * 16; field = 1; break;
        case FHTGS_DBG_VPMT4: src = sl.vb.prim; bytes = N * 16; field = 2; break;
        case FHTGS_DBG_RGBA: src = sl.vb.prim; bytes = N * 16; field = 3; break;
        case FHTGS_DBG_KEYS_UNSORTED:
            if (!h->keep_unsorted) return set_err(h, FHTGS_E_INVALID, "FHTGS_OPT_KEEP_UNSORTED is off");
            src = sl.dbg_keys; bytes = I * 4; break;
        case FHTGS_DBG_VALS_UNSORTED:
            if (!h->keep_unsorted) return set_err(h, FHTGS_E_INVALID, "FHTGS_OPT_KEEP_UNSORTED is off");
            src = sl.dbg_vals; bytes = I * 4; break;
        case FHTGS_DBG_KEYS_SORTED: src = sl.ib.keys[sl.sorted_src]; bytes = I * 4; widen = sl.plan.key_bytes == 2; break;
        case FHTGS_DBG_VALS_SORTED: src = sl.ib.vals[sl.sorted_src]; bytes = I * 4; break;
        case FHTGS_DBG_RANGES: src = sl.vb.ranges; bytes = (int64_t)sl.vp.n_small * 8; break;
        default: return set_err(h, FHTGS_E_INVALID, "unknown debug buffer id");
    }
    if (bytes > dst_bytes) return set_err(h, FHTGS_E_INVALID, "destination too small");
    int prev = 0;
    HCHK(hipGetDevice(&prev));
    HCHK(hipSetDevice(h->device));
    if (h->have_done) HCHK(hipEventSynchronize(h->done));
    if (bytes > 0) {
        if (widen) {
            std::vector<uint16_t> tmp((size_t)I);
            HCHK(hipMemcpy(tmp.data(), src, (size_t)I * 2, hipMemcpyDeviceToHost));
            uint32_t* o = (uint32_t*)dst;
            for (int64_t k = 0; k < I; ++k) o[k] = tmp[(size_t)k];
        } else if (field >= 0) {
            // one float4 out of each 64-B primitive record: whole records in
            // contiguous chunks, the field picked out on the host (a pitched
            // copy of millions of 16-B rows is slow)
            const int64_t chunk = (int64_t)1 << 20;
            std::vector<float4> tmp((size_t)std::min(N, chunk) * 4);
            const float4* rec = (const float4*)src;
            char* o = (char*)dst;
            for (int64_t r0 = 0; r0 < N; r0 += chunk) {
                const int64_t nr = std::min(chunk, N - r0);
                HCHK(hipMemcpy(tmp.data(), rec + 4 * r0, (size_t)nr * 64, hipMemcpyDeviceToHost));
                for (int64_t k = 0; k < nr; ++k)
                    std::memcpy(o + (size_t)(r0 + k) * 16, &tmp[(size_t)(4 * k + field)], 16);
            }
        } else {
            HCHK(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToHost));
        }
    }
    HCHK(hipSetDevice(prev));
    if (written) *written = bytes;
    return FHTGS_OK;
}

namespace {

// ---- scratch of fhtgs_sort_pairs_noalloc: one fixed array per device ----
// Alternate keys and values (4-byte keys at most), digit histograms,
// tickets plus the error word, and the look-back status of up to kMaxPasses
// passes of 8-bit digits over kNoAllocMaxPairs pairs, each part 256-B aligned.
constexpr int64_t kNoAllocMaxPairs = FHTGS_SORT_NOALLOC_MAX_PAIRS;
constexpr int kNoAllocDigitBits = 8;
constexpr size_t a256(size_t x) { return (x + 255) & ~(size_t)255; }
constexpr int64_t kNoAllocTiles = (kNoAllocMaxPairs + kSortTile - 1) / kSortTile;
constexpr size_t kNoAllocOffVals = a256(4 * (size_t)kNoAllocMaxPairs);
constexpr size_t kNoAllocOffHist = kNoAllocOffVals + a256(4 * (size_t)kNoAllocMaxPairs);
constexpr size_t kNoAllocOffTickets = kNoAllocOffHist + a256(4 * (size_t)kMaxPasses * kMaxBins);
constexpr size_t kNoAllocOffStatus = kNoAllocOffTickets + a256(4 * (size_t)(kMaxPasses + 1));
constexpr size_t kNoAllocBytes =
    kNoAllocOffStatus + a256(4 * (size_t)kMaxPasses * kNoAllocTiles * (1u << kNoAllocDigitBits));
static_assert(kNoAllocBytes < (size_t)16 << 20, "the fixed sort scratch stays modest");

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

// Device address of `scratch` per device, from hipGetSymbolAddress on the
// first call on that device (the array's name is not a device address on
// the host). Every call on a device shares that device's copy.
constexpr int kNoAllocMaxDevices = 64;
std::mutex g_scratch_lock;
void* g_scratch_ptr[kNoAllocMaxDevices] = {};

int scratch_for_device(int dev, void** out) {
    if (dev < 0 || dev >= kNoAllocMaxDevices)
        return set_err(nullptr, FHTGS_E_UNSUPPORTED, "device index beyond the fixed-scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_lock);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_err(nullptr, FHTGS_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return FHTGS_OK;
}

}  // namespace

int fhtgs_sort_pairs_noalloc(void* keys, uint32_t* values, int64_t n, int32_t key_bytes, int32_t end_bit,
                             uint32_t* error, void* stream_ptr) {
    fhtgs_context* h = nullptr;  // no handle: messages go to fhtgs_last_error(NULL)
    if (n < 0 || n > kNoAllocMaxPairs)
        return set_err(h, FHTGS_E_INVALID,
                       "n out of range: the fixed scratch holds at most " + std::to_string(kNoAllocMaxPairs) + " pairs");
    if (key_bytes != 2 && key_bytes != 4) return set_err(h, FHTGS_E_INVALID, "key_bytes must be 2 or 4");
    if (end_bit < 1 || end_bit > key_bytes * 8) return set_err(h, FHTGS_E_INVALID, "end_bit out of range");
    if (n > 0 && (!keys || !values)) return set_err(h, FHTGS_E_INVALID, "null keys or values");
    hipStream_t s = (hipStream_t)stream_ptr;
    int dev = 0;
    HCHK(hipGetDevice(&dev));
    void* base = nullptr;
    int rc = scratch_for_device(dev, &base);
    if (rc) return rc;
    SortPlan plan = make_plan(end_bit, kNoAllocDigitBits);
    plan.key_bytes = key_bytes;
    const int64_t tiles = (n + kSortTile - 1) / kSortTile;
    InstanceBuffers ib;
    char* b = (char*)base;
    ib.keys[0] = keys;
    ib.vals[0] = values;
    ib.keys[1] = b;
    ib.vals[1] = (uint32_t*)(b + kNoAllocOffVals);
    ib.hist = (uint32_t*)(b + kNoAllocOffHist);
    ib.tickets = (uint32_t*)(b + kNoAllocOffTickets);
    ib.status = (uint32_t*)(b + kNoAllocOffStatus);
    ib.capacity = n;
    ib.tiles_cap = tiles;
    // histograms, tickets, error word and status start at zero (a fill
    // kernel, so the call can be captured into a graph)
    HCHK(launch_sort_clear(ib, plan, tiles, s));
    if (n > 0) {
        HCHK(launch_sort_histogram(ib, plan, 0, (uint32_t)n, s));
        int src = 0;
        for (int p = 0; p < plan.passes; ++p) {
            HCHK(launch_sort_pass(ib, plan, p, src, tiles, nullptr, (uint32_t)n, s));
            src = 1 - src;
        }
        if (src == 1) {
            HCHK(hipMemcpyAsync(keys, ib.keys[1], (size_t)n * key_bytes, hipMemcpyDeviceToDevice, s));
            HCHK(hipMemcpyAsync(values, ib.vals[1], (size_t)n * 4, hipMemcpyDeviceToDevice, s));
        }
    }
    if (error) HCHK(hipMemcpyAsync(error, ib.tickets + kMaxPasses, sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    return FHTGS_OK;
}

int fhtgs_sort_pairs(fhtgs_handle h, void* keys, uint32_t* values, int64_t n, int32_t key_bytes,
                     int32_t end_bit, void* stream_ptr) {
    if (!h) return FHTGS_E_INVALID;
    if (n < 0 || n >= ((int64_t)1 << 30)) return set_err(h, FHTGS_E_INVALID, "n out of range");
    if (key_bytes != 2 && key_bytes != 4) return set_err(h, FHTGS_E_INVALID, "key_bytes must be 2 or 4");
    if (end_bit < 1 || end_bit > key_bytes * 8) return set_err(h, FHTGS_E_INVALID, "end_bit out of range");
    if (n == 0) return FHTGS_OK;
    if (!keys || !values) return set_err(h, FHTGS_E_INVALID, "null keys or values");
    hipStream_t s = (hipStream_t)stream_ptr;
    int prev = 0;
    HCHK(hipGetDevice(&prev));
    if (prev != h->device) HCHK(hipSetDevice(h->device));
    if (h->have_done) HCHK(hipStreamWaitEvent(s, h->done, 0));
    SortPlan plan = make_plan(end_bit, h->digit_bits);
    plan.key_bytes = key_bytes;
    // scratch: alternate keys/values, histograms, tickets, status
    InstanceBuffers ib;
    size_t bytes = 0;
    carve_inst(nullptr, n, plan, ib, bytes);
    int rc = ensure(h, h->sort_scratch, bytes);
    if (rc) return rc;
    carve_inst(h->sort_scratch.ptr, n, plan, ib, bytes);
    void* alt_keys = ib.keys[1];
    uint32_t* alt_vals = ib.vals[1];
    ib.keys[0] = keys;
    ib.vals[0] = values;
    ib.keys[1] = alt_keys;
    ib.vals[1] = alt_vals;
    const int64_t tiles = (n + kSortTile - 1) / kSortTile;
    HCHK(launch_sort_clear(ib, plan, tiles, s));
    HCHK(launch_sort_histogram(ib, plan, 0, (uint32_t)n, s));
    int src = 0;
    for (int p = 0; p < plan.passes; ++p) {
        HCHK(launch_sort_pass(ib, plan, p, src, tiles, nullptr, (uint32_t)n, s));
        src = 1 - src;
    }
    if (src == 1) {
        HCHK(hipMemcpyAsync(keys, alt_keys, (size_t)n * key_bytes, hipMemcpyDeviceToDevice, s));
        HCHK(hipMemcpyAsync(values, alt_vals, (size_t)n * 4, hipMemcpyDeviceToDevice, s));
    }
    HCHK(hipEventRecord(h->done, s));
    h->have_done = true;
    HCHK(hipStreamSynchronize(s));
    uint32_t sort_err = 0;
    HCHK(hipMemcpy(&sort_err, ib.tickets + kMaxPasses, sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (prev != h->device) HCHK(hipSetDevice(prev));
    if (sort_err) return set_err(h, FHTGS_E_HIP, "sort detected an inconsistency (look-back or scatter)");
    return FHTGS_OK;
}

}  // extern "C"
