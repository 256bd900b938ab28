// detect.hpp -- device pieces of the single-maximum detector shared by the
// kernels that search windows for their best S/N (search_kernels.hip and
// the allocation-free variant): the candidate record, its total order, the
// per-cell window scan and the final reduction over per-block bests.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"

namespace sps {
namespace {

struct Best {
    float snr;
    int32_t w, r, t;
};

// a total order (NaN values never win), so the best of any partition of the
// windows is the same window
__device__ __forceinline__ bool better(const Best& a, const Best& b) {
    if (a.w < 0 || !(a.snr == a.snr)) return false;
    if (b.w < 0 || !(b.snr == b.snr)) return true;
    if (a.snr != b.snr) return a.snr > b.snr;
    if (a.w != b.w) return a.w < b.w;
    if (a.r != b.r) return a.r < b.r;
    return a.t < b.t;
}

__device__ void block_best(Best v, Best* sh) {
    sh[threadIdx.x] = v;
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        __syncthreads();
        if ((int)threadIdx.x < s && better(sh[threadIdx.x + s], sh[threadIdx.x])) sh[threadIdx.x] = sh[threadIdx.x + s];
    }
    __syncthreads();
}

// columns of one detector tile (one thread per starting column)
constexpr int kDetCells = 256;

// every window of 1..wnd samples starting at (r, t), of x / sqrtf(1e-7 + n)
// summed left to right, divided by sqrtf(w), into the running best b
__device__ __forceinline__ void detect_windows(const float* __restrict__ a, const float* __restrict__ n, int64_t r,
                                               int64_t t, int64_t cols, int wnd, Best& b) {
    float s = 0.0f;
    for (int w = 1; w <= wnd && t + w <= cols; ++w) {
        float v = a[r * cols + t + w - 1];
        if (n) v = v / sqrtf((float)(1.0E-7 + (double)n[r * cols + t + w - 1]));
        s += v;
        const float val = s / sqrtf((float)w);
        Best c{val, w, (int32_t)r, (int32_t)t};
        if (better(c, b)) b = c;
    }
}

// one block: the best of nb per-block results into *out
__global__ void __launch_bounds__(kThreads) k_detect_final(const Best* __restrict__ blocks, int64_t nb,
                                                           sps_detection* __restrict__ out) {
    __shared__ Best sh[kThreads];
    Best b{0.0f, -1, -1, -1};
    for (int64_t i = threadIdx.x; i < nb; i += kThreads)
        if (better(blocks[i], b)) b = blocks[i];
    block_best(b, sh);
    if (threadIdx.x == 0) {
        out->iwidth = sh[0].w;
        out->irow = sh[0].r;
        out->icol = sh[0].t;
        out->snr = sh[0].snr;
    }
}

}  // namespace
}  // namespace sps
