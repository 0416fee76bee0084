// Device-side vector helpers for the gfx950 kernels. Float arithmetic follows
// the convention of fhtgs_math.h: no implicit contraction; each "a*b + c" of
// the reference expressions is one fmaf; in "a*b - c*d" the first product is
// fused; dot products are left-to-right fmaf chains.
#pragma once

#include <hip/hip_runtime.h>
#include "fhtgs_math.h"
#include "fhtgs_internal.h"

namespace fhtgs {

struct v3 { float x, y, z; };
struct v4 { float x, y, z, w; };

__device__ __forceinline__ v4 ld4(const float4& f) { return v4{f.x, f.y, f.z, f.w}; }
__device__ __forceinline__ float4 st4(const v4& v) { return make_float4(v.x, v.y, v.z, v.w); }

// helper_math.h:1245-1252
__device__ __forceinline__ float dot3(v3 a, v3 b) {
    return fmaf(a.z, b.z, fmaf(a.y, b.y, a.x * b.x));
}
__device__ __forceinline__ float dot3(v4 a, v3 b) {
    return fmaf(a.z, b.z, fmaf(a.y, b.y, a.x * b.x));
}
__device__ __forceinline__ float dot4(v4 a, v4 b) {
    return fmaf(a.w, b.w, fmaf(a.z, b.z, fmaf(a.y, b.y, a.x * b.x)));
}
__device__ __forceinline__ v4 mul4(v4 a, v4 b) { return v4{a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w}; }
// helper_math.h:1429-1432
__device__ __forceinline__ v3 cross3(v3 a, v3 b) {
    return v3{fmaf(a.y, b.z, -(a.z * b.y)), fmaf(a.z, b.x, -(a.x * b.z)), fmaf(a.x, b.y, -(a.y * b.x))};
}
// helper_math.h:1306-1310, rsqrtf taken as 1/sqrtf
__device__ __forceinline__ v3 normalize3(v3 v) {
    const float inv = 1.0f / sqrtf(dot3(v, v));
    return v3{v.x * inv, v.y * inv, v.z * inv};
}

// Correctly rounded 1/x: v_rcp_f32 plus one FMA Newton step equals the IEEE
// quotient for every float with 2^-125 <= |x| <= 2^125 on gfx950 (checked
// exhaustively by rcp_check, built next to the library); other x take the
// IEEE division.
// fh_expf_m64 with its last step in gfx950 form, bit-identical on [-64, 0]
// and NaN (csrc/rcp_check.hip compares the two on every such input): the
// scaling by 2^k is v_ldexp_f32, exact for these normal results as the
// multiply by the power of two is; for NaN x, p is NaN and so is the result,
// whatever v_cvt_i32_f32 gives for k (the portable form's select is not
// needed).
__device__ __forceinline__ float fh_expf_m64_gfx(float x) {
    float k;
    const float p = fh_expf_core(x, &k);
    int32_t ki;
    __asm__("v_cvt_i32_f32 %0, %1" : "=v"(ki) : "v"(k));
    return __builtin_amdgcn_ldexpf(p, ki);
}

__device__ __forceinline__ float rcp_rn(float x) {
    const float r0 = __builtin_amdgcn_rcpf(x);
    float r = fmaf(fmaf(-x, r0, 1.0f), r0, r0);
    const float ax = fabsf(x);
    const bool fast = ax >= 0x1p-125f && ax <= 0x1p125f;
    // the IEEE division only when some lane needs it (a wave-uniform branch,
    // so it is not evaluated speculatively)
    if (__builtin_amdgcn_ballot_w64(!fast) != 0ull) r = fast ? r : 1.0f / x;
    return r;
}

// Pointers read from the device parameter block are generic to the
// compiler, which then emits flat loads and stores through them. Reading
// each pointer field through a global-address-space view of its storage
// makes the loaded value a global pointer, and the accesses global ones.
// Every such pointer is a hipMalloc'd device address.
template <typename T>
__device__ __forceinline__ T* load_gptr(T* const* field) {
    typedef __attribute__((address_space(1))) T GT;
    return (T*)(*(GT* const*)field);
}

// Kernel-side copy of a view's parameters with global pointers.
__device__ __forceinline__ ViewParams load_view(const ViewParams* src) {
    ViewParams v = *src;
    v.M = load_gptr(&src->M);
    v.VPM = load_gptr(&src->VPM);
    v.VPR_inv = load_gptr(&src->VPR_inv);
    v.cam = load_gptr(&src->cam);
    v.vis_mask = load_gptr(&src->vis_mask);
    v.vis_sat = load_gptr(&src->vis_sat);
    v.fovea_sat = load_gptr(&src->fovea_sat);
    v.bg_data = load_gptr(&src->bg_data);
    v.image_final = load_gptr(&src->image_final);
    v.image_blend = load_gptr(&src->image_blend);
    return v;
}

__device__ __forceinline__ v4 load_row(const float* m, int row) {
    return v4{m[4 * row + 0], m[4 * row + 1], m[4 * row + 2], m[4 * row + 3]};
}

// kernel_utils.cuh:43-73
__device__ __forceinline__ bool in_fovea(int radius, float cx, float cy, float gx, float gy) {
    const float radius_sq = (float)(radius * radius);
    const float dx = cx - gx, dy = cy - gy;
    return fmaf(dy, dy, dx * dx) < radius_sq;
}

__device__ __forceinline__ bool mask_bit(const uint32_t* mask, uint32_t t) {
    return (mask[t >> 5] & (1u << (t & 31u))) != 0u;
}

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t lanemask_lt() {
    return (1ull << lane_id()) - 1ull;
}

// Wave64 inclusive prefix sum.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= (uint32_t)o) x += y;
    }
    return x;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

__device__ __forceinline__ uint32_t atomic_load_agent(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void atomic_store_agent(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace fhtgs
