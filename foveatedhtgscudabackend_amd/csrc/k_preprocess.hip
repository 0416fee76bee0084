// Preprocess: one thread per Gaussian, any number of views (1 = mono, 2 =
// stereo with one shared read of the scene). Per view: view-depth cull, the
// rows VPMT1/2/4 of VPM*T, the exact screen-space extent of the quadric at
// the opacity cutoff, tile bounds, tiles-touched from the two summed-area
// tables, and the degree-0..3 SH colour. Also writes the per-block sum of
// tiles-touched, which the instance kernel turns into write offsets, and
// each visible Gaussian's support-ellipse record (fh_ellipse_record), which
// the blend's tile cull and first pass read instead of deriving it again for
// every tile the Gaussian is listed in.
//
// Follows rasterization/include/kernels/inference.cuh:42-136,
// rasterization/include/utils/kernel_utils.cuh:76-213 and
// rasterization/include/utils/sh_utils.cuh:32-69.
#include "fhtgs_device.h"
#include "fhtgs_cull.h"

namespace fhtgs {

constexpr int kMaxViewsPerPass = 4;
// LDS slots of SH coefficients per wave (a visible Gaussian each; a wave
// with more visible Gaussians copies the rest in a further round). 44 slots
// of 45 floats are 31.7 KB per workgroup of four waves: the most that still
// lets the five workgroups per CU which the registers allow share the LDS.
constexpr int kShSlots = 44;

// Kernel argument: the NV views' parameters and the scene's tensors by value
// (pointers read from kernel arguments are global ones to the compiler). The
// preprocess therefore does not wait for the frame's parameter upload, which
// only the later kernels read.
template <int NV>
struct PreprocessArgs {
    ViewParams vp[NV];
    ViewBuffers vb[NV];
    SceneArgs scene;
    uint32_t n;
};

// The same, resolved into registers at kernel start.
template <int NV>
struct PreprocessState {
    ViewParams vp[NV];
    ViewBuffers vb[NV];
    const float* pos;
    const float* scl;
    const float* rot;
    const float* opa;
    const float* sh0;
    const float* shr;
    uint32_t n;
};

template <int NV>
__device__ __forceinline__ PreprocessState<NV> resolve(const PreprocessArgs<NV>& g) {
    PreprocessState<NV> a;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        a.vp[v] = g.vp[v];
        a.vb[v] = g.vb[v];
    }
    const SceneArgs& sc = g.scene;
    a.pos = sc.pos; a.scl = sc.scl; a.rot = sc.rot; a.opa = sc.opa; a.sh0 = sc.sh0; a.shr = sc.shr;
    a.n = g.n;
    return a;
}

// sh_utils.cuh:9-30
constexpr float C0 = 0.28209479177387814f;
constexpr float C1 = 0.48860251190291987f;
constexpr float C2a = 1.0925484305920792f;
constexpr float C2b = 0.94617469575755997f;
constexpr float C2c = 0.31539156525251999f;
constexpr float C2d = 0.54627421529603959f;
constexpr float C3a = 0.59004358992664352f;
constexpr float C3b = 1.7701307697799304f;
constexpr float C3c = 2.8906114426405538f;
constexpr float C3d = 0.45704579946446572f;
constexpr float C3e = 2.2852289973223288f;
constexpr float C3f = 1.865881662950577f;
constexpr float C3g = 1.1195289977703462f;
constexpr float C3h = 1.4453057213202769f;

__device__ __forceinline__ void sh_acc(v3& r, float t, const float* k) {
    r.x = fmaf(t, k[0], r.x);
    r.y = fmaf(t, k[1], r.y);
    r.z = fmaf(t, k[2], r.z);
}

// sh_utils.cuh:32-69
__device__ __forceinline__ v3 sh_color(const float* sh0, const float* shr, v3 mu, v3 cam,
                                       int active, int total) {
    v3 r{fmaf(C0, sh0[0], 0.5f), fmaf(C0, sh0[1], 0.5f), fmaf(C0, sh0[2], 0.5f)};
    if (active > 1) {
        const v3 n = normalize3(v3{mu.x - cam.x, mu.y - cam.y, mu.z - cam.z});
        const float x = n.x, y = n.y, z = n.z;
        sh_acc(r, -(C1 * y), shr + 0);
        sh_acc(r, C1 * z, shr + 3);
        sh_acc(r, -(C1 * x), shr + 6);
        if (active > 4) {
            const float xx = x * x, yy = y * y, zz = z * z;
            const float xy = x * y, xz = x * z, yz = y * z;
            sh_acc(r, C2a * xy, shr + 9);
            sh_acc(r, -(C2a * yz), shr + 12);
            sh_acc(r, fmaf(C2b, zz, -C2c), shr + 15);
            sh_acc(r, -(C2a * xz), shr + 18);
            sh_acc(r, C2d * (xx - yy), shr + 21);
            if (active > 9) {
                sh_acc(r, y * fmaf(C3a, yy, -(C3b * xx)), shr + 24);
                sh_acc(r, (C3c * xy) * z, shr + 27);
                sh_acc(r, y * fmaf(-C3e, zz, C3d), shr + 30);
                sh_acc(r, z * fmaf(C3f, zz, -C3g), shr + 33);
                sh_acc(r, x * fmaf(-C3e, zz, C3d), shr + 36);
                sh_acc(r, (C3h * z) * (xx - yy), shr + 39);
                sh_acc(r, x * fmaf(C3b, yy, -(C3a * xx)), shr + 42);
            }
        }
    }
    return r;
}

// Row `row` of VPM times [u v w | mu] (kernel_utils.cuh:135,144,152,156).
__device__ __forceinline__ v4 frame_row(v4 row, v3 u, v3 v, v3 w, v3 mu) {
    return v4{dot3(row, u), dot3(row, v), dot3(row, w), dot3(row, mu) + row.w};
}

// the view's record row and support-ellipse record of a visible Gaussian
// (the ellipse as 32 B, so that the blend gathers it in one memory sector;
// 20 B in two arrays cost the blend a second sector per listed Gaussian). The
// logarithm is taken again here rather than kept in a register through the
// SH gathers, which need every register they can get.
__device__ __forceinline__ void store_record(const ViewBuffers& vb, uint32_t i, v4 V1, v4 V2, v4 V4,
                                             float4 rgba) {
    const float c2 = 2.0f * fh_logf(rgba.w * 255.0f);
    float4* rec = vb.prim + 4 * (size_t)i;
    rec[0] = st4(V1);
    rec[1] = st4(V2);
    rec[2] = st4(V4);
    rec[3] = rgba;
    const float a1[4] = {V1.x, V1.y, V1.z, V1.w}, a2[4] = {V2.x, V2.y, V2.z, V2.w};
    const float a4[4] = {V4.x, V4.y, V4.z, V4.w};
    float r[5];
    fh_ellipse_record(a1, a2, a4, c2, r);
    vb.ell[2 * (size_t)i] = make_float4(r[0], r[1], r[2], r[3]);
    vb.ell[2 * (size_t)i + 1] = make_float4(r[4], 0.0f, 0.0f, 0.0f);
}

__device__ __forceinline__ uint32_t tile_lo(float c, float e, int grid) {
    const int t = max(0, fh_f2i_rd(((c - e) + 0.5f) / 16.0f));
    return min((uint32_t)grid, (uint32_t)t);
}
__device__ __forceinline__ uint32_t tile_hi(float c, float e, int grid) {
    const int t = max(0, fh_f2i_ru(((c + e) + 0.5f) / 16.0f));
    return min((uint32_t)grid, (uint32_t)t);
}

// ---- anti-aliased variant (rasterization/include/utils/kernel_utils_aaa.cuh)

__device__ __forceinline__ v4 row_minus(v4 a, v4 b, float s) {
    return v4{fmaf(-b.x, s, a.x), fmaf(-b.y, s, a.y), fmaf(-b.z, s, a.z), fmaf(-b.w, s, a.w)};
}

// kernel_utils_aaa.cuh:9-25
__device__ __forceinline__ float aaa_contrib_plane(v4 pl, v4 V1, v4 V2, v4 V3, v4 V4, v4& ps) {
    const float norm = pl.w / fmaf(pl.z, pl.z, fmaf(pl.y, pl.y, pl.x * pl.x));
    const v3 g{(-norm) * pl.x, (-norm) * pl.y, (-norm) * pl.z};
    const float gw = (-norm) * pl.w;
    ps = v4{dot3(V1, g) + V1.w, dot3(V2, g) + V2.w, dot3(V3, g) + V3.w, dot3(V4, g) + V4.w};
    return -gw;
}

// kernel_utils_aaa.cuh:27-65
__device__ __forceinline__ float aaa_contrib_ray(v4 a, v4 b, v4 V1, v4 V2, v4 V3, v4 V4, v4& ps) {
    const v3 d = cross3(v3{a.x, a.y, a.z}, v3{b.x, b.y, b.z});
    const v3 m{fmaf(a.w, b.x, -(a.x * b.w)), fmaf(a.w, b.y, -(a.y * b.w)), fmaf(a.w, b.z, -(a.z * b.w))};
    const float dd = dot3(d, d);
    const v3 mdd{m.x / dd, m.y / dd, m.z / dd};
    const v3 mp = cross3(d, mdd);
    ps = v4{dot3(V1, mp) + V1.w, dot3(V2, mp) + V2.w, dot3(V3, mp) + V3.w, dot3(V4, mp) + V4.w};
    return dot3(m, mdd);
}

__device__ __forceinline__ bool aaa_in_range(float p, float from, float ext, float ph) {
    const float d = fmaf(-from, ph, p);
    return d > 0.0f && d < ext * ph;
}

// kernel_utils_aaa.cuh:77-133
__device__ float aaa_contrib_frustum(v4 V1, v4 V2, v4 V3, v4 V4, float width, float height) {
    float best = 3.402823466e+38f;
    const float rf0 = 0.0f, rf1 = 0.0f, rf2 = -1.0f;
    const float ex0 = width - 1.0f, ex1 = height - 1.0f, ex2 = 2.0f;
    const float d0 = fmaf(-rf0, V4.w, V1.w), d1 = fmaf(-rf1, V4.w, V2.w), d2 = fmaf(-rf2, V4.w, V3.w);
    const bool bx = d0 > 0.0f && d0 < ex0 * V4.w;
    const bool by = d1 > 0.0f && d1 < ex1 * V4.w;
    const bool bz = d2 > 0.0f && d2 < ex2 * V4.w;
    if (bx && by && bz) return 0.0f;
    const float dx = copysignf(ex0 * 0.5f, fmaf(-(ex0 * 0.5f), V4.w, d0));
    const float dy = copysignf(ex1 * 0.5f, fmaf(-(ex1 * 0.5f), V4.w, d1));
    const v4 cpx = row_minus(V1, V4, fmaf(ex0, 0.5f, rf0) + dx);
    const v4 cpy = row_minus(V2, V4, fmaf(ex1, 0.5f, rf1) + dy);
    v4 ps;
    float c = aaa_contrib_plane(cpx, V1, V2, V3, V4, ps);
    if (c < best && aaa_in_range(ps.y, rf1, ex1, ps.w) && aaa_in_range(ps.z, rf2, ex2, ps.w)) best = c;
    c = aaa_contrib_plane(cpy, V1, V2, V3, V4, ps);
    if (c < best && aaa_in_range(ps.x, rf0, ex0, ps.w) && aaa_in_range(ps.z, rf2, ex2, ps.w)) best = c;
    c = aaa_contrib_ray(cpx, cpy, V1, V2, V3, V4, ps);
    if (c < best && aaa_in_range(ps.z, rf2, ex2, ps.w)) best = c;
    const v4 opy = row_minus(V2, V4, fmaf(ex1, 0.5f, rf1) - dy);
    c = aaa_contrib_ray(cpx, opy, V1, V2, V3, V4, ps);
    if (c < best && aaa_in_range(ps.z, rf2, ex2, ps.w)) best = c;
    const v4 opx = row_minus(V1, V4, fmaf(ex0, 0.5f, rf0) - dx);
    c = aaa_contrib_ray(opx, cpy, V1, V2, V3, V4, ps);
    if (c < best && aaa_in_range(ps.z, rf2, ex2, ps.w)) best = c;
    return best;
}

// kernel_utils_aaa.cuh:135-144 (fmodf is exact)
__device__ __forceinline__ float aaa_normalize_angle(float th) {
    const float pi = 3.141592654f;
    const float two_pi = 2.0f * pi;
    th = fmodf(th, two_pi);
    if (th > pi) th -= two_pi;
    if (th <= -pi) th += two_pi;
    return th;
}

// kernel_utils_aaa.cuh:162-201, one axis: tan of the clamped angle range
__device__ void aaa_theta(v3 vd, v4 MTa, v4 MT3, float cutoff, bool is_x, float& lo, float& hi) {
    const float pi = 3.141592654f;
    const float two_pi = 2.0f * pi;
    const float max_theta = fmaf(0.5f, pi, -1e-5f);
    const v4 t{cutoff, cutoff, cutoff, -1.0f};
    const float mu = fh_atan2f(is_x ? vd.x : vd.y, vd.z);
    const float sa = dot4(t, mul4(MTa, MTa)), sz = dot4(t, mul4(MT3, MT3)), mid = dot4(t, mul4(MTa, MT3));
    const float inside = fmaf(mid, mid, -(sz * sa));
    float r0 = -max_theta, r1 = max_theta;
    if (inside > 0.0f) {
        const float sq = sqrtf(inside);
        float a0 = fh_atan2f(-(mid + sq), -sz);
        float a1 = fh_atan2f(-(mid - sq), -sz);
        while (a0 > mu) a0 -= pi;
        while (a0 < (mu - pi)) a0 += pi;
        while (a1 < mu) a1 += pi;
        while (a1 > (mu + pi)) a1 -= pi;
        const float n0 = aaa_normalize_angle(a0), n1 = aaa_normalize_angle(a1);
        if (mu < 0.0f && fabsf(n0) < fabsf(n1)) { a0 += two_pi; a1 += two_pi; }
        else if (mu > 0.0f && fabsf(n1) < fabsf(n0)) { a0 -= two_pi; a1 -= two_pi; }
        r0 = fmaxf(r0, a0);
        r1 = fminf(r1, a1);
    }
    lo = fh_tanf(r0);
    hi = fh_tanf(r1);
}

struct Frame {
    float R11, R12, R13, R21, R22, R23, R31, R32, R33;
};

// Culled or not; outputs as the non-AAA path plus the dilated opacity
// (kernel_utils_aaa.cuh:210-363).
__device__ bool aaa_transform_cull(const ViewParams& vp, const Frame& F, v3 mu, v3 os, float& opacity,
                                   uint32_t& b0, uint32_t& b1, uint32_t& b2, uint32_t& b3, v4& V1,
                                   v4& V2, v4& V4) {
    const v4 M3 = load_row(vp.M, 2);
    const float z = dot3(M3, mu) + M3.w;
    const float max_focal = fmaxf(vp.fx, vp.fy);
    const float fsv = z / max_focal;
    const float fvv = (fsv * fsv) * 0.3f;
    const v3 ov{os.x * os.x, os.y * os.y, os.z * os.z};
    const v3 var{ov.x + fvv, ov.y + fvv, ov.z + fvv};
    const v3 s{sqrtf(var.x), sqrtf(var.y), sqrtf(var.z)};
    const v3 cam{vp.cam[0], vp.cam[1], vp.cam[2]};
    const v3 dw = normalize3(v3{mu.x - cam.x, mu.y - cam.y, mu.z - cam.z});
    const v3 dg{fmaf(F.R31, dw.z, fmaf(F.R21, dw.y, F.R11 * dw.x)), fmaf(F.R32, dw.z, fmaf(F.R22, dw.y, F.R12 * dw.x)),
                fmaf(F.R33, dw.z, fmaf(F.R23, dw.y, F.R13 * dw.x))};
    const v3 rr{dg.x * dg.x, dg.y * dg.y, dg.z * dg.z};
    const float det0 = dot3(rr, v3{ov.y * ov.z, ov.z * ov.x, ov.x * ov.y});
    const float det1 = dot3(rr, v3{var.y * var.z, var.z * var.x, var.x * var.y});
    opacity = opacity * sqrtf(fh_saturate(det0 / det1));
    if (opacity < kMinAlpha) return true;
    const float rho = 2.0f * fh_logf(opacity * kMinAlphaRcp);
    const v3 cs{cam.x - mu.x, cam.y - mu.y, cam.z - mu.z};
    const v3 cg{fmaf(F.R31, cs.z, fmaf(F.R21, cs.y, F.R11 * cs.x)) / s.x,
                fmaf(F.R32, cs.z, fmaf(F.R22, cs.y, F.R12 * cs.x)) / s.y,
                fmaf(F.R33, cs.z, fmaf(F.R23, cs.y, F.R13 * cs.x)) / s.z};
    if (dot3(cg, cg) < rho) return true;
    const float sm = vp.scale_mod;
    const v3 u{(F.R11 * s.x) * sm, (F.R21 * s.x) * sm, (F.R31 * s.x) * sm};
    const v3 vv{(F.R12 * s.y) * sm, (F.R22 * s.y) * sm, (F.R32 * s.y) * sm};
    const v3 w{(F.R13 * s.z) * sm, (F.R23 * s.z) * sm, (F.R33 * s.z) * sm};
    V1 = frame_row(load_row(vp.VPM, 0), u, vv, w, mu);
    V2 = frame_row(load_row(vp.VPM, 1), u, vv, w, mu);
    const v4 V3 = frame_row(load_row(vp.VPM, 2), u, vv, w, mu);
    V4 = frame_row(load_row(vp.VPM, 3), u, vv, w, mu);
    if (aaa_contrib_frustum(V1, V2, V3, V4, (float)vp.width, (float)vp.height) > rho) return true;
    const v4 MT1 = frame_row(load_row(vp.M, 0), u, vv, w, mu);
    const v4 MT2 = frame_row(load_row(vp.M, 1), u, vv, w, mu);
    const v3 vd = normalize3(v3{MT1.w, MT2.w, V4.w});
    float tx0, tx1, ty0, ty1;
    aaa_theta(vd, MT1, V4, rho, true, tx0, tx1);
    aaa_theta(vd, MT2, V4, rho, false, ty0, ty1);
    const float bx0 = fmaf(vp.fx, tx0, vp.cx), bx1 = fmaf(vp.fx, tx1, vp.cx);
    const float by0 = fmaf(vp.fy, ty0, vp.cy), by1 = fmaf(vp.fy, ty1, vp.cy);
    const float ccx = (bx1 + bx0) / 2.0f, ccy = (by1 + by0) / 2.0f;
    const float ecx = (bx1 - bx0) / 2.0f, ecy = (by1 - by0) / 2.0f;
    // kernel_utils_aaa.cuh:311-316: no half-pixel shift in this variant
    b0 = min((uint32_t)vp.gw, (uint32_t)max(0, fh_f2i_rd((ccx - ecx) / 16.0f)));
    b1 = min((uint32_t)vp.gw, (uint32_t)max(0, fh_f2i_ru((ccx + ecx) / 16.0f)));
    b2 = min((uint32_t)vp.gh, (uint32_t)max(0, fh_f2i_rd((ccy - ecy) / 16.0f)));
    b3 = min((uint32_t)vp.gh, (uint32_t)max(0, fh_f2i_ru((ccy + ecy) / 16.0f)));
    return false;
}

// The most instances any Gaussian can have: kNumSub per tile of the largest
// grid. A summed-area table of another image size or radius can give a
// rectangle sum below zero, which wraps to a value far above this; such a
// value would wrap the instance offsets and is refused. (Smaller wrong
// counts only under- or over-fill the Gaussian's own range, which the
// instance kernel reports or fills with sentinels.)
constexpr uint32_t kMaxTouched = (uint32_t)kNumSub * (uint32_t)kMaxLargeTiles;
constexpr uint32_t kTouchedRefused = 0xffffffffu;

// Tiles-touched from the visibility and fovea summed-area tables
// (kernel_utils.cuh:168-210): at most kNumSub per tile of the bounds when
// the tables fit the view; the callers refuse values above kMaxTouched.
__device__ __forceinline__ uint32_t tiles_touched(const ViewParams& vp, uint32_t b0, uint32_t b1, uint32_t b2,
                                                  uint32_t b3) {
    const int fd = 2 * vp.fovea_r;
    const int tlx = vp.gtx - vp.fovea_r, tly = vp.gty - vp.fovea_r;
    const uint32_t f0 = (uint32_t)min(fd, max(0, (int)b0 - tlx));
    const uint32_t f1 = (uint32_t)min(fd, max(0, (int)b1 - tlx));
    const uint32_t f2 = (uint32_t)min(fd, max(0, (int)b2 - tly));
    const uint32_t f3 = (uint32_t)min(fd, max(0, (int)b3 - tly));
    const uint32_t W1 = (uint32_t)vp.gw + 1u;
    const uint32_t* vs = vp.vis_sat;
    uint32_t nt = vs[b1 + W1 * b3] + vs[b0 + W1 * b2] - vs[b1 + W1 * b2] - vs[b0 + W1 * b3];
    if (f1 != f0 && f3 != f2) {
        const uint32_t F1 = (uint32_t)fd + 1u;
        const uint32_t* fs = vp.fovea_sat;
        nt += fs[f1 + F1 * f3] + fs[f0 + F1 * f2] - fs[f1 + F1 * f2] - fs[f0 + F1 * f3];
    }
    return nt;
}

template <int NV>
__global__ __launch_bounds__(kPrimBlock) void preprocess_aaa_kernel(PreprocessArgs<NV> g) {
    const PreprocessState<NV> a = resolve<NV>(g);
    __shared__ uint32_t s_sum[NV][kPrimBlock / 64];
    __shared__ uint32_t s_vis[NV][kPrimBlock / 64];
    const uint32_t i = blockIdx.x * kPrimBlock + threadIdx.x;
    const bool valid = i < a.n;
    uint32_t touched[NV];
    bool visible[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) { touched[v] = 0u; visible[v] = false; }
    if (valid) {
        const v3 mu{a.pos[3 * i + 0], a.pos[3 * i + 1], a.pos[3 * i + 2]};
        const v3 os{a.scl[3 * i + 0], a.scl[3 * i + 1], a.scl[3 * i + 2]};
        const float4 q = *reinterpret_cast<const float4*>(a.rot + 4 * (size_t)i);
        const float r = q.x, x = q.y, y = q.z, z = q.w;
        const float xx = x * x, yy = y * y, zz = z * z;
        const float xy = x * y, xz = x * z, yz = y * z;
        const float rx = r * x, ry = r * y, rz = r * z;
        const Frame F{fmaf(-2.0f, yy + zz, 1.0f), 2.0f * (xy - rz), 2.0f * (xz + ry),
                      2.0f * (xy + rz), fmaf(-2.0f, xx + zz, 1.0f), 2.0f * (yz - rx),
                      2.0f * (xz - ry), 2.0f * (yz + rx), fmaf(-2.0f, xx + yy, 1.0f)};
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const ViewParams& vp = a.vp[v];
            float opacity = a.opa[i];
            uint32_t b0, b1, b2, b3;
            v4 V1, V2, V4;
            if (aaa_transform_cull(vp, F, mu, os, opacity, b0, b1, b2, b3, V1, V2, V4)) continue;
            const uint32_t nt = tiles_touched(vp, b0, b1, b2, b3);
            if (nt > kMaxTouched) {
                // tables that do not fit the view: marked in the word that is
                // live anyway, reported and cleared at the end; the Gaussian
                // reserves nothing
                touched[v] = kTouchedRefused;
                continue;
            }
            if (nt == 0u) continue;
            touched[v] = nt;
            visible[v] = true;
            const ViewBuffers& vb = a.vb[v];
            vb.bounds[i] = make_ushort4((unsigned short)b0, (unsigned short)b1, (unsigned short)b2, (unsigned short)b3);
            const v3 cam{vp.cam[0], vp.cam[1], vp.cam[2]};
            const v3 rgb = sh_color(a.sh0 + 3 * (size_t)i, a.shr + (size_t)i * vp.total_sh * 3, mu, cam,
                                    vp.active_sh, vp.total_sh);
            store_record(vb, i, V1, V2, V4, make_float4(rgb.x, rgb.y, rgb.z, opacity));
        }
    }
    const uint32_t lane = lane_id(), wid = threadIdx.x >> 6;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        const bool refused = touched[v] == kTouchedRefused;
        if (refused) touched[v] = 0u;
        const uint32_t refused_w = __ballot(refused) != 0ull ? kVisRefused : 0u;
        if (valid) a.vb[v].n_touched[i] = touched[v];
        const uint32_t ws = wave_sum(touched[v]);
        const uint32_t wv = (uint32_t)__popcll(__ballot(visible[v])) | refused_w;
        if (lane == 0) { s_sum[v][wid] = ws; s_vis[v][wid] = wv; }
    }
    __syncthreads();
    // thread v finishes view v (static indices keep the parameters in registers)
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        if ((int)threadIdx.x != v) continue;
        uint32_t tot = 0, vis = 0;
        uint32_t flag = 0;
        for (int w = 0; w < kPrimBlock / 64; ++w) {
            tot += s_sum[v][w];
            vis += s_vis[v][w] & ~kVisRefused;
            flag |= s_vis[v][w] & kVisRefused;
        }
        // per block, so that the frame-info words need no clearing before
        // the preprocess: the block scan sums the counts and flags
        a.vb[v].block_sums[blockIdx.x] = tot;
        a.vb[v].block_vis[blockIdx.x] = vis | flag;
    }
}

template <int NV>
__global__ __launch_bounds__(kPrimBlock) void preprocess_kernel(PreprocessArgs<NV> g) {
    const PreprocessState<NV> a = resolve<NV>(g);
    extern __shared__ float s_sh[];  // per wave: kShSlots slots of the first min(3S, 45) SH floats
    __shared__ uint32_t s_sum[NV][kPrimBlock / 64];
    __shared__ uint32_t s_vis[NV][kPrimBlock / 64];
    const uint32_t i = blockIdx.x * kPrimBlock + threadIdx.x;
    const bool valid = i < a.n;
    const uint32_t lane = lane_id(), wid = threadIdx.x >> 6;

    uint32_t touched[NV];
    bool visible[NV];
    // per view: the record rows and bounds, stored after every load (one
    // counter tracks loads and stores, so loads after stores wait for them)
    v4 R1[NV], R2[NV], R4[NV];
    ushort4 bnd[NV];
    float4 col[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) { touched[v] = 0u; visible[v] = false; }
    v3 mu{0.0f, 0.0f, 0.0f};
    float opacity = 0.0f, s0[3] = {0.0f, 0.0f, 0.0f};
    bool zok[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) zok[v] = false;
    v3 sc{0.0f, 0.0f, 0.0f};
    float R11 = 0.0f, R12 = 0.0f, R13 = 0.0f, R21 = 0.0f, R22 = 0.0f, R23 = 0.0f, R31 = 0.0f, R32 = 0.0f,
          R33 = 0.0f, rho = 0.0f;
    if (valid) {
        mu = v3{a.pos[3 * i + 0], a.pos[3 * i + 1], a.pos[3 * i + 2]};
        opacity = a.opa[i];
        s0[0] = a.sh0[3 * (size_t)i + 0];
        s0[1] = a.sh0[3 * (size_t)i + 1];
        s0[2] = a.sh0[3 * (size_t)i + 2];
        // view depth cull first (kernel_utils.cuh:121-123)
        bool any = false;
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const v4 M3 = load_row(a.vp[v].M, 2);
            const float z = dot3(M3, mu) + M3.w;
            zok[v] = !(z < a.vp[v].near_plane || z > a.vp[v].far_plane);
            any |= zok[v];
        }
        if (any) {
            sc = v3{a.scl[3 * i + 0], a.scl[3 * i + 1], a.scl[3 * i + 2]};
            const float4 q = *reinterpret_cast<const float4*>(a.rot + 4 * (size_t)i);
            // kernel_utils.cuh:76-88, quaternion (r, x, y, z)
            const float r = q.x, x = q.y, y = q.z, z = q.w;
            const float xx = x * x, yy = y * y, zz = z * z;
            const float xy = x * y, xz = x * z, yz = y * z;
            const float rx = r * x, ry = r * y, rz = r * z;
            R11 = fmaf(-2.0f, yy + zz, 1.0f); R12 = 2.0f * (xy - rz); R13 = 2.0f * (xz + ry);
            R21 = 2.0f * (xy + rz); R22 = fmaf(-2.0f, xx + zz, 1.0f); R23 = 2.0f * (yz - rx);
            R31 = 2.0f * (xz - ry); R32 = 2.0f * (yz + rx); R33 = fmaf(-2.0f, xx + yy, 1.0f);
            rho = 2.0f * fh_logf(opacity * kMinAlphaRcp);
        }
    }
    // one view's transform, extent, tile bounds and tiles-touched
    // (kernel_utils.cuh:76-213)
    auto transform = [&](int v) {
        if (!zok[v]) return;
        const ViewParams& vp = a.vp[v];
        const float sm = vp.scale_mod;
        const v3 u{(R11 * sc.x) * sm, (R21 * sc.x) * sm, (R31 * sc.x) * sm};
        const v3 vv{(R12 * sc.y) * sm, (R22 * sc.y) * sm, (R32 * sc.y) * sm};
        const v3 w{(R13 * sc.z) * sm, (R23 * sc.z) * sm, (R33 * sc.z) * sm};
        const v4 V4 = frame_row(load_row(vp.VPM, 3), u, vv, w, mu);
        const v4 d{rho, rho, rho, -1.0f};
        const float s = dot4(d, mul4(V4, V4));
        if (s == 0.0f) return;
        const float rs = 1.0f / s;
        const v4 f{rs * d.x, rs * d.y, rs * d.z, rs * d.w};
        const v4 V3 = frame_row(load_row(vp.VPM, 2), u, vv, w, mu);
        const float cz = dot4(f, mul4(V3, V4));
        const float ez = sqrtf(fmaxf(fmaf(cz, cz, -dot4(f, mul4(V3, V3))), 0.0f));
        if (cz - ez < -1.0f || cz + ez > 1.0f) return;
        const v4 V1 = frame_row(load_row(vp.VPM, 0), u, vv, w, mu);
        const float cx = dot4(f, mul4(V1, V4));
        const float ex = sqrtf(fmaxf(fmaf(cx, cx, -dot4(f, mul4(V1, V1))), 0.0f));
        const v4 V2 = frame_row(load_row(vp.VPM, 1), u, vv, w, mu);
        const float cy = dot4(f, mul4(V2, V4));
        const float ey = sqrtf(fmaxf(fmaf(cy, cy, -dot4(f, mul4(V2, V2))), 0.0f));
        const uint32_t b0 = tile_lo(cx, ex, vp.gw), b1 = tile_hi(cx, ex, vp.gw);
        const uint32_t b2 = tile_lo(cy, ey, vp.gh), b3 = tile_hi(cy, ey, vp.gh);
        const uint32_t nt = tiles_touched(vp, b0, b1, b2, b3);
        if (nt > kMaxTouched) {
            // tables that do not fit the view: marked in the word that is
            // live anyway, reported and cleared at the end; the Gaussian
            // reserves nothing
            touched[v] = kTouchedRefused;
            return;
        }
        if (nt == 0u) return;
        touched[v] = nt;
        visible[v] = true;
        bnd[v] = make_ushort4((unsigned short)b0, (unsigned short)b1, (unsigned short)b2, (unsigned short)b3);
        R1[v] = V1;
        R2[v] = V2;
        R4[v] = V4;
    };

    // SH colour of the visible Gaussians (sh_utils.cuh:32-69). The wave
    // copies the coefficients of each visible Gaussian into an LDS slot of
    // its own with one direct-to-LDS load across the lanes (global_load_lds:
    // no registers held, all in flight): those of the Gaussians visible in
    // the first view right after that view's transform, so that they arrive
    // while the second view's transform runs, the rest after it. A wave with
    // more visible Gaussians than slots copies the remainder in a further
    // round. The owning lanes then evaluate from LDS.
    const int active = a.vp[0].active_sh;
    const int total = a.vp[0].total_sh;
    const int shw = min(3 * total, 45);
    const bool gather = active > 1;  // degree 0 needs no coefficients
    float* wsh = s_sh + (size_t)wid * kShSlots * shw;
    uint32_t used = 0;  // slots taken (wave-uniform)
    int my_slot = -1;
    // lanes of m take the next free slots in lane order; returns those left over
    auto issue = [&](uint64_t m) -> uint64_t {
        while (m != 0ull && used < (uint32_t)kShSlots) {
            const int src = __ffsll((unsigned long long)m) - 1;
            m &= m - 1ull;
            const uint32_t ip = __builtin_amdgcn_readlane(i, src);
            if ((int)lane == src) my_slot = (int)used;
            if ((int)lane < shw)
                __builtin_amdgcn_global_load_lds(a.shr + (size_t)ip * (3 * total) + lane,
                                                 (__attribute__((address_space(3))) void*)(wsh + used * shw), 4, 0, 0);
            ++used;
        }
        return m;
    };
    uint64_t rest = 0ull;
    transform(0);
    if (gather) rest = issue(__ballot(visible[0]));
    if constexpr (NV == 2) {
        transform(1);
        if (gather) rest |= issue(__ballot(visible[1] && !visible[0]));
    } else if constexpr (NV > 2) {
        bool seen = visible[0];  // visible in a view before v: its coefficients are on their way
#pragma unroll
        for (int v = 1; v < NV; ++v) {
            transform(v);
            if (gather) rest |= issue(__ballot(visible[v] && !seen));
            seen |= visible[v];
        }
    }
    // evaluates the lanes holding a slot, then frees the slots
    auto evaluate = [&]() {
        __builtin_amdgcn_s_waitcnt(0);  // the direct-to-LDS loads have landed
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (my_slot >= 0) {
            const float* coef = wsh + my_slot * shw;
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                if (!visible[v]) continue;
                const ViewParams& vp = a.vp[v];
                const v3 cam{vp.cam[0], vp.cam[1], vp.cam[2]};
                const v3 rgb = sh_color(s0, coef, mu, cam, vp.active_sh, vp.total_sh);
                col[v] = make_float4(rgb.x, rgb.y, rgb.z, opacity);
            }
        }
        // every LDS read of the slots has returned before a further round's
        // direct-to-LDS loads may overwrite them
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        my_slot = -1;
        used = 0;
    };
    if (gather) {
        evaluate();
        while (rest != 0ull) {  // more visible Gaussians than slots (rare)
            rest = issue(rest);
            evaluate();
        }
    } else {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            if (!visible[v]) continue;
            const ViewParams& vp = a.vp[v];
            const v3 cam{vp.cam[0], vp.cam[1], vp.cam[2]};
            const v3 rgb = sh_color(s0, a.shr, mu, cam, vp.active_sh, vp.total_sh);
            col[v] = make_float4(rgb.x, rgb.y, rgb.z, opacity);
        }
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        if (!visible[v]) continue;
        const ViewBuffers& vb = a.vb[v];
        vb.bounds[i] = bnd[v];
        store_record(vb, i, R1[v], R2[v], R4[v], col[v]);
    }

#pragma unroll
    for (int v = 0; v < NV; ++v) {
        const bool refused = touched[v] == kTouchedRefused;
        if (refused) touched[v] = 0u;
        const uint32_t refused_w = __ballot(refused) != 0ull ? kVisRefused : 0u;
        if (valid) a.vb[v].n_touched[i] = touched[v];
        const uint32_t ws = wave_sum(touched[v]);
        const uint32_t wv = (uint32_t)__popcll(__ballot(visible[v])) | refused_w;
        if (lane == 0) { s_sum[v][wid] = ws; s_vis[v][wid] = wv; }
    }
    __syncthreads();
    // thread v finishes view v (static indices keep the parameters in registers)
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        if ((int)threadIdx.x != v) continue;
        uint32_t tot = 0, vis = 0;
        uint32_t flag = 0;
        for (int w = 0; w < kPrimBlock / 64; ++w) {
            tot += s_sum[v][w];
            vis += s_vis[v][w] & ~kVisRefused;
            flag |= s_vis[v][w] & kVisRefused;
        }
        // per block, so that the frame-info words need no clearing before
        // the preprocess: the block scan sums the counts and flags
        a.vb[v].block_sums[blockIdx.x] = tot;
        a.vb[v].block_vis[blockIdx.x] = vis | flag;
    }
}

template <int NV>
hipError_t launch_preprocess_nv(const ViewParams* const* vps, const ViewBuffers* vbs, const SceneArgs& scene,
                                uint32_t n, hipStream_t s) {
    PreprocessArgs<NV> a;
    for (int v = 0; v < NV; ++v) {
        a.vp[v] = *vps[v];
        a.vb[v] = vbs[v];
    }
    a.scene = scene;
    a.n = n;
    const uint32_t blocks = (n + kPrimBlock - 1) / kPrimBlock;
    // inference.cuh:86-112: the anti-aliased variant replaces transform and cull
    if (vps[0]->aaa) {
        preprocess_aaa_kernel<NV><<<blocks, kPrimBlock, 0, s>>>(a);
    } else {
        const size_t lds = sizeof(float) * (size_t)(kPrimBlock / 64) * kShSlots * (size_t)min(3 * vps[0]->total_sh, 45);
        preprocess_kernel<NV><<<blocks, kPrimBlock, lds, s>>>(a);
    }
    return hipGetLastError();
}

hipError_t launch_preprocess(const ViewParams* const* vps, const ViewBuffers* vbs, int n_views,
                             const SceneArgs& scene, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    switch (n_views) {
        case 1: return launch_preprocess_nv<1>(vps, vbs, scene, n, s);
        case 2: return launch_preprocess_nv<2>(vps, vbs, scene, n, s);
        case 3: return launch_preprocess_nv<3>(vps, vbs, scene, n, s);
        case 4: return launch_preprocess_nv<4>(vps, vbs, scene, n, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace fhtgs
