// Checks that fh_expf_nonpos (the blend's range-specialised exp) equals
// fh_expf bit for bit on [-64, 0] (every `stride`-th float) and on NaN.
#include <cstdio>
#include <cstdlib>
#include "../foveatedhtgscudabackend_amd/csrc/fhtgs_math.h"

int main(int argc, char** argv) {
    const uint32_t stride = argc > 1 ? (uint32_t)std::atoi(argv[1]) : 1u;
    long bad = 0, n = 0;
    for (uint32_t u = 0x80000000u;; u += stride) {
        const float x = fh_as_f32(u);
        if (x < -64.0f) break;
        const float a = fh_expf(x), b = fh_expf_nonpos(x);
        if (fh_as_u32(a) != fh_as_u32(b)) {
            if (bad < 5) std::printf("x=%a fh_expf=%a fast=%a\n", x, a, b);
            ++bad;
        }
        ++n;
    }
    const float qnan = fh_as_f32(0x7fc00000u);
    if (!(fh_expf_nonpos(qnan) != fh_expf_nonpos(qnan))) ++bad;
    std::printf("checked %ld, mismatches %ld\n", n, bad);
    return bad != 0;
}
