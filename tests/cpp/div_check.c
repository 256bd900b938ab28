/* Bit-for-bit check of the shared-reciprocal division used by the kernels
 * (div_rn in cpp_projects_amd/csrc/common.hpp): with r = RN(1/y),
 * q0 = x*r, e = fma(-q0, y, x), q1 = fma(e, r, q0) must equal RN(x/y).
 * Every x mantissa in [1, 2) is tried (other binades scale exactly) against
 * every odd integer divisor up to argv[1] (powers of two scale exactly) and
 * argv[2] pseudo-random float divisors over 2^-20..2^20, and sqrtf(2) (the
 * boxcar's odd-width divisors are it times powers of two).  Prints the number
 * of mismatches. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static long check(float fy) {
    const float r = 1.0f / fy;
    long bad = 0;
    for (uint32_t m = 0; m < (1u << 23); ++m) {
        const uint32_t bits = 0x3f800000u | m;
        float x;
        memcpy(&x, &bits, 4);
        volatile float q = x / fy;
        const float q0 = x * r;
        const float e = fmaf(-q0, fy, x);
        const float q1 = fmaf(e, r, q0);
        const float qq = q;
        uint32_t a, b;
        memcpy(&a, &qq, 4);
        memcpy(&b, &q1, 4);
        bad += a != b;
    }
    return bad;
}

int main(int argc, char** argv) {
    const int ymax = argc > 1 ? atoi(argv[1]) : 511;
    const int nrand = argc > 2 ? atoi(argv[2]) : 64;
    long bad = 0;
#pragma omp parallel for schedule(dynamic) reduction(+ : bad)
    for (int k = 0; k < (ymax + 1) / 2 + nrand; ++k) {
        float fy;
        if (k < (ymax + 1) / 2) {
            fy = (float)(2 * k + 1);
        } else {
            uint64_t s = 0x9E3779B97F4A7C15ull * (uint64_t)(k + 1);
            s ^= s >> 31;
            s *= 0xBF58476D1CE4E5B9ull;
            s ^= s >> 29;
            const uint32_t yb = ((uint32_t)(127 + (int)((s >> 32) % 41) - 20) << 23) | (uint32_t)(s & 0x7fffff);
            memcpy(&fy, &yb, 4);
        }
        bad += check(fy);
    }
    /* the boxcar's odd-width divisor sqrtf(2^p) = sqrtf(2) * 2^((p-1)/2) */
    bad += check(sqrtf(2.0f));
    printf("%ld\n", bad);
    return 0;
}
