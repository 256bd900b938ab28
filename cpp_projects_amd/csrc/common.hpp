// common.hpp -- shared host/device helpers for libsps.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

#include "../../include/sps.h"

namespace sps {

struct Error : std::runtime_error {
    sps_status status;
    Error(sps_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};

#define SPS_HIP(expr)                                                                                   \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess)                                                                           \
            throw ::sps::Error(e_ == hipErrorOutOfMemory ? SPS_ERR_NO_MEMORY : SPS_ERR_HIP,              \
                               std::string(#expr) + ": " + hipGetErrorString(e_));                      \
    } while (0)

#define SPS_REQUIRE(cond, msg)                                                          \
    do {                                                                                \
        if (!(cond)) throw ::sps::Error(SPS_ERR_INVALID_ARGUMENT, std::string(msg));    \
    } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Device buffer owned by the host object that holds it.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    void ensure(size_t n) {
        if (n <= bytes && p) return;
        release();
        if (n == 0) n = 256;
        SPS_HIP(hipMalloc(&p, n));
        bytes = n;
    }
    template <class T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

__host__ __device__ inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Correctly rounded x / y given r = 1.0f / y (the reciprocal rounded to
// nearest), for divisors shared by many quotients: q0 = x * r is within one
// ulp of x / y, the FMA residual x - q0 * y is exact, and RN(q0 + r * residual)
// is the correctly rounded quotient (Markstein, IBM J. Res. Dev. 34(1), 1990).
// tests/test_division.py checks it bit for bit against x / y over every x
// mantissa for integer divisors and random float divisors.  Operands outside
// [2^-60, 2^60] in magnitude (zero, tiny, huge, infinite, NaN) take the
// division, so no intermediate under- or overflows.
__device__ __forceinline__ float div_rn(float x, float y, float r) {
    const float ax = fabsf(x), ay = fabsf(y);
    if (ax >= 0x1p-60f && ax <= 0x1p60f && ay >= 0x1p-60f && ay <= 0x1p60f) {
        const float q0 = x * r;
        const float e = fmaf(-q0, y, x);
        return fmaf(e, r, q0);
    }
    return x / y;
}

// Launch geometry: the merge/init kernels tile rows of up to this many samples.
constexpr int kThreads = 256;

}  // namespace sps
