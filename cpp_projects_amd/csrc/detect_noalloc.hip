// detect_noalloc.hip -- the single-maximum detector for callers that cannot
// allocate device memory (code under stream capture, tight serving loops,
// harnesses that forbid hipMalloc): sps_detect_max_noalloc gives the result
// of sps_detect_max with its per-block partial results in a static device
// array instead of a buffer allocated per call.
#include <map>
#include <mutex>

#include "abi.hpp"
#include "common.hpp"
#include "detect.hpp"

// Per-block results: kScratchBlocks entries of 16 bytes (64 KiB), one copy
// per GPU, held until the process ends.  Host code never uses this name as
// an address: it uses the device address hipGetSymbolAddress returns for
// the current device (scratch_ptr()).
namespace {
constexpr int kScratchBlocks = 4096;
}
static __device__ __align__(256) unsigned char scratch[kScratchBlocks * 16];

namespace sps {
namespace {

// The windows of sps_detect_max over a fixed number of blocks: block k takes
// the (row, 256-column tile) pairs k, k + gridDim.x, ... and writes one best,
// so the per-block results fit kScratchBlocks entries whatever the image size.
__global__ void __launch_bounds__(kThreads) k_detect_blocks_strided(const float* __restrict__ a,
                                                                    const float* __restrict__ n, int64_t rows,
                                                                    int64_t cols, int wnd,
                                                                    Best* __restrict__ blocks) {
    __shared__ Best sh[kThreads];
    const int64_t per_row = (cols + kDetCells - 1) / kDetCells;
    const int64_t ntiles = per_row * rows;
    Best b{0.0f, -1, -1, -1};
    for (int64_t k = blockIdx.x; k < ntiles; k += gridDim.x) {
        const int64_t r = k / per_row;
        const int64_t t = (k - r * per_row) * kDetCells + threadIdx.x;
        if (t < cols) detect_windows(a, n, r, t, cols, wnd, b);
    }
    block_best(b, sh);
    if (threadIdx.x == 0) blocks[blockIdx.x] = sh[0];
}

// The device address of `scratch` on the current device, looked up on the
// first call made with that device current and kept per device.
void* scratch_ptr() {
    static std::mutex mu;
    static std::map<int, void*> addr;
    int dev = 0;
    SPS_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    auto it = addr.find(dev);
    if (it != addr.end()) return it->second;
    void* p = nullptr;
    SPS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw Error(SPS_ERR_HIP, "no device address for the detector scratch");
    addr[dev] = p;
    return p;
}

}  // namespace
}  // namespace sps

using namespace sps;

extern "C" {

SPS_API sps_status sps_detect_max_noalloc(const float* d_arr, const float* d_norm, int64_t rows, int64_t cols, int wnd,
                                          sps_detection* d_out, void* stream) {
    return guarded([&] {
        SPS_REQUIRE(d_arr && d_out && rows > 0 && cols > 0 && wnd >= 1, "invalid argument");
        SPS_REQUIRE(rows <= 0x7fffffffLL && cols <= 0x7fffffffLL,
                    "rows and columns must fit 32 bits (the result holds them as int)");
        static_assert(sizeof(Best) == 16, "one scratch entry per block");
        const int64_t ntiles = ceil_div(cols, kDetCells) * rows;
        const int64_t nb = ntiles < kScratchBlocks ? ntiles : kScratchBlocks;
        SPS_REQUIRE(nb >= 1 && nb <= kScratchBlocks, "more detector blocks than the scratch holds");
        Best* blocks = static_cast<Best*>(scratch_ptr());
        hipStream_t st = as_stream(stream);
        // every entry the final reduction reads was written by its block
        // earlier on the same stream, so nothing needs clearing
        hipLaunchKernelGGL(k_detect_blocks_strided, dim3((uint32_t)nb), dim3(kThreads), 0, st, d_arr, d_norm, rows,
                           cols, wnd, blocks);
        hipLaunchKernelGGL(k_detect_final, dim3(1), dim3(kThreads), 0, st, static_cast<const Best*>(blocks), nb, d_out);
        SPS_HIP(hipGetLastError());
    });
}

}  // extern "C"
