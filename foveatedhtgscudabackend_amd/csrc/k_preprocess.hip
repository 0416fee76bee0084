// Preprocess: one thread per Gaussian, any number of views (1 = mono, 2 =
// stereo with one shared read of the scene). Per view: view-depth cull, the
// rows VPMT1/2/4 of VPM*T, the exact screen-space extent of the quadric at
// the opacity cutoff, tile bounds, tiles-touched from the two summed-area
// tables, and the degree-0..3 SH colour. Also writes the per-block sum of
// tiles-touched, which the instance kernel turns into write offsets.
//
// Follows rasterization/include/kernels/inference.cuh:42-136,
// rasterization/include/utils/kernel_utils.cuh:76-213 and
// rasterization/include/utils/sh_utils.cuh:32-69.
#include "fhtgs_device.h"

namespace fhtgs {

constexpr int kMaxViewsPerPass = 2;

struct PreprocessArgs {
    ViewParams vp[kMaxViewsPerPass];
    ViewBuffers vb[kMaxViewsPerPass];
    const float* pos;
    const float* scl;
    const float* rot;
    const float* opa;
    const float* sh0;
    const float* shr;
    uint32_t n;
};

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

__device__ __forceinline__ uint32_t tile_lo(float c, float e, int grid) {
    const int t = max(0, fh_f2i_rd(((c - e) + 0.5f) / 16.0f));
    return min((uint32_t)grid, (uint32_t)t);
}
__device__ __forceinline__ uint32_t tile_hi(float c, float e, int grid) {
    const int t = max(0, fh_f2i_ru(((c + e) + 0.5f) / 16.0f));
    return min((uint32_t)grid, (uint32_t)t);
}

template <int NV>
__global__ __launch_bounds__(kPrimBlock) void preprocess_kernel(PreprocessArgs a) {
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
        const float opacity = a.opa[i];
        // view depth cull first (kernel_utils.cuh:121-123)
        bool any = false;
        bool zok[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const v4 M3 = load_row(a.vp[v].M, 2);
            const float z = dot3(M3, mu) + M3.w;
            zok[v] = !(z < a.vp[v].near_plane || z > a.vp[v].far_plane);
            any |= zok[v];
        }
        if (any) {
            const v3 sc{a.scl[3 * i + 0], a.scl[3 * i + 1], a.scl[3 * i + 2]};
            const float4 q = *reinterpret_cast<const float4*>(a.rot + 4 * (size_t)i);
            // kernel_utils.cuh:76-88, quaternion (r, x, y, z)
            const float r = q.x, x = q.y, y = q.z, z = q.w;
            const float xx = x * x, yy = y * y, zz = z * z;
            const float xy = x * y, xz = x * z, yz = y * z;
            const float rx = r * x, ry = r * y, rz = r * z;
            const float R11 = fmaf(-2.0f, yy + zz, 1.0f), R12 = 2.0f * (xy - rz), R13 = 2.0f * (xz + ry);
            const float R21 = 2.0f * (xy + rz), R22 = fmaf(-2.0f, xx + zz, 1.0f), R23 = 2.0f * (yz - rx);
            const float R31 = 2.0f * (xz - ry), R32 = 2.0f * (yz + rx), R33 = fmaf(-2.0f, xx + yy, 1.0f);
            const float rho = 2.0f * fh_logf(opacity * kMinAlphaRcp);
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                if (!zok[v]) continue;
                const ViewParams& vp = a.vp[v];
                const float sm = vp.scale_mod;
                const v3 u{(R11 * sc.x) * sm, (R21 * sc.x) * sm, (R31 * sc.x) * sm};
                const v3 vv{(R12 * sc.y) * sm, (R22 * sc.y) * sm, (R32 * sc.y) * sm};
                const v3 w{(R13 * sc.z) * sm, (R23 * sc.z) * sm, (R33 * sc.z) * sm};
                const v4 V4 = frame_row(load_row(vp.VPM, 3), u, vv, w, mu);
                const v4 d{rho, rho, rho, -1.0f};
                const float s = dot4(d, mul4(V4, V4));
                if (s == 0.0f) continue;
                const float rs = 1.0f / s;
                const v4 f{rs * d.x, rs * d.y, rs * d.z, rs * d.w};
                const v4 V3 = frame_row(load_row(vp.VPM, 2), u, vv, w, mu);
                const float cz = dot4(f, mul4(V3, V4));
                const float ez = sqrtf(fmaxf(fmaf(cz, cz, -dot4(f, mul4(V3, V3))), 0.0f));
                if (cz - ez < -1.0f || cz + ez > 1.0f) continue;
                const v4 V1 = frame_row(load_row(vp.VPM, 0), u, vv, w, mu);
                const float cx = dot4(f, mul4(V1, V4));
                const float ex = sqrtf(fmaxf(fmaf(cx, cx, -dot4(f, mul4(V1, V1))), 0.0f));
                const v4 V2 = frame_row(load_row(vp.VPM, 1), u, vv, w, mu);
                const float cy = dot4(f, mul4(V2, V4));
                const float ey = sqrtf(fmaxf(fmaf(cy, cy, -dot4(f, mul4(V2, V2))), 0.0f));
                const uint32_t b0 = tile_lo(cx, ex, vp.gw), b1 = tile_hi(cx, ex, vp.gw);
                const uint32_t b2 = tile_lo(cy, ey, vp.gh), b3 = tile_hi(cy, ey, vp.gh);
                // kernel_utils.cuh:168-210
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
                if (nt == 0u) continue;
                touched[v] = nt;
                visible[v] = true;
                const ViewBuffers& vb = a.vb[v];
                vb.bounds[i] = make_ushort4((unsigned short)b0, (unsigned short)b1, (unsigned short)b2,
                                            (unsigned short)b3);
                vb.vpmt1[i] = st4(V1);
                vb.vpmt2[i] = st4(V2);
                vb.vpmt4[i] = st4(V4);
                const v3 cam{vp.cam[0], vp.cam[1], vp.cam[2]};
                const v3 rgb = sh_color(a.sh0 + 3 * (size_t)i, a.shr + (size_t)i * vp.total_sh * 3, mu,
                                        cam, vp.active_sh, vp.total_sh);
                vb.rgba[i] = make_float4(rgb.x, rgb.y, rgb.z, opacity);
            }
        }
    }

    const uint32_t lane = lane_id(), wid = threadIdx.x >> 6;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        if (valid) a.vb[v].n_touched[i] = touched[v];
        const uint32_t ws = wave_sum(touched[v]);
        const uint32_t wv = (uint32_t)__popcll(__ballot(visible[v]));
        if (lane == 0) { s_sum[v][wid] = ws; s_vis[v][wid] = wv; }
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        const int v = threadIdx.x;
        uint32_t tot = 0, vis = 0;
        for (int w = 0; w < kPrimBlock / 64; ++w) { tot += s_sum[v][w]; vis += s_vis[v][w]; }
        a.vb[v].block_sums[blockIdx.x] = tot;
        if (vis) atomicAdd(&a.vb[v].frame_info[FI_N_VISIBLE], vis);
    }
}

hipError_t launch_preprocess(const ViewParams* vps, const ViewBuffers* vbs, int n_views,
                             const float* pos, const float* scl, const float* rot,
                             const float* opa, const float* sh0, const float* shr, uint32_t n,
                             int, hipStream_t s) {
    if (n == 0) return hipSuccess;
    if (n_views < 1 || n_views > kMaxViewsPerPass) return hipErrorInvalidValue;
    PreprocessArgs a;
    for (int v = 0; v < n_views; ++v) { a.vp[v] = vps[v]; a.vb[v] = vbs[v]; }
    a.pos = pos; a.scl = scl; a.rot = rot; a.opa = opa; a.sh0 = sh0; a.shr = shr; a.n = n;
    const uint32_t blocks = (n + kPrimBlock - 1) / kPrimBlock;
    if (n_views == 1) preprocess_kernel<1><<<blocks, kPrimBlock, 0, s>>>(a);
    else preprocess_kernel<2><<<blocks, kPrimBlock, 0, s>>>(a);
    return hipGetLastError();
}

}  // namespace fhtgs
