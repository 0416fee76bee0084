// rcp_check: exhaustive checks of the blend's two gfx950 fast paths
// (fhtgs_device.h): rcp_rn returns the IEEE-754 quotient 1/x bit for bit for
// every one of the 2^32 float bit patterns, and fh_expf_m64_gfx equals the
// portable fh_expf_m64 bit for bit on every float of [-64, 0] (NaN results
// compared as NaN). Prints the number of mismatches and the first few; exit
// status 0 iff there are none.
#include "fhtgs_device.h"

#include <cstdio>
#include <cstring>

namespace {

__global__ __launch_bounds__(256) void rcp_check_kernel(unsigned long long* count, uint32_t* first) {
    const uint64_t total = 1ull << 32;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const float x = __uint_as_float((uint32_t)i);
        const float want = 1.0f / x;
        const float got = fhtgs::rcp_rn(x);
        const bool both_nan = want != want && got != got;
        bool bad = !both_nan && __float_as_uint(want) != __float_as_uint(got);
        // the exp fast path on [-64, 0] (including -0) and every NaN
        if ((x <= 0.0f && x >= -64.0f) || x != x) {
            const float e0 = fh_expf_m64(x), e1 = fhtgs::fh_expf_m64_gfx(x);
            const bool e_nan = e0 != e0 && e1 != e1;
            bad |= !e_nan && __float_as_uint(e0) != __float_as_uint(e1);
        }
        if (bad) {
            const unsigned long long k = atomicAdd(count, 1ull);
            if (k < 8ull) first[k] = (uint32_t)i;
        }
    }
}

}  // namespace

int main() {
    unsigned long long* d_count = nullptr;
    uint32_t* d_first = nullptr;
    if (hipMalloc(&d_count, sizeof(unsigned long long)) != hipSuccess ||
        hipMalloc(&d_first, 8 * sizeof(uint32_t)) != hipSuccess ||
        hipMemset(d_count, 0, sizeof(unsigned long long)) != hipSuccess ||
        hipMemset(d_first, 0, 8 * sizeof(uint32_t)) != hipSuccess) {
        std::printf("rcp_check: HIP set-up failed\n");
        return 2;
    }
    rcp_check_kernel<<<32768, 256>>>(d_count, d_first);
    unsigned long long count = 0;
    uint32_t first[8] = {};
    if (hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(&count, d_count, sizeof(count), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(first, d_first, sizeof(first), hipMemcpyDeviceToHost) != hipSuccess) {
        std::printf("rcp_check: kernel failed\n");
        return 2;
    }
    std::printf("rcp_check: 4294967296 inputs, mismatches %llu\n", count);
    for (unsigned long long k = 0; k < count && k < 8ull; ++k) {
        float x;
        std::memcpy(&x, &first[k], 4);
        std::printf("  x = %a (0x%08x)\n", x, first[k]);
    }
    if (hipFree(d_count) != hipSuccess || hipFree(d_first) != hipSuccess) {
        std::printf("rcp_check: hipFree failed\n");
        return 2;
    }
    return count == 0 ? 0 : 1;
}
