// fdmt_kernels.hip -- FDMT init, merge and adjoint kernels for gfx950.
//
// Layout: every state is row-major [rows][pitch] fp32, rows time-contiguous.
// A merge row reads two rows of the previous state (tail at t, head at
// t - offset) and writes one; rows are independent, so one launch covers an
// iteration.  Blocks are mapped XCD-aware: the hardware deals consecutive
// block ids round-robin to the 8 XCDs, so the id is remapped to give each XCD
// one contiguous range of (time tile, row) work, row fastest.  Rows that
// share an input row (neighbouring dt of one sub-band) then run at the same
// time on the same XCD and the shared row is served from that XCD's L2.
#include "common.hpp"
#include "kernels.hpp"

namespace sps {

namespace {

constexpr int kMergeVecPerThread = 4;                             // float4 per thread
constexpr int kMergeTile = kThreads * 4 * kMergeVecPerThread;     // 4096 samples per block
constexpr int kNumXcd = 8;

__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t nblocks) {
    const uint32_t xcd = b % kNumXcd, k = b / kNumXcd;
    const uint32_t q = nblocks / kNumXcd, rem = nblocks % kNumXcd;
    return xcd * q + min(xcd, rem) + k;
}

typedef float float4u __attribute__((ext_vector_type(4), aligned(4)));
typedef float float4a __attribute__((ext_vector_type(4), aligned(16)));

// ---------------------------------------------------------------------------
// merge
// ---------------------------------------------------------------------------
template <bool VEC>
__global__ void __launch_bounds__(kThreads) k_merge(const float* __restrict__ in, int64_t in_pitch,
                                                    float* __restrict__ out, int64_t out_pitch,
                                                    const RowDescD* __restrict__ rows, int64_t nrows,
                                                    int64_t nsamps, int64_t out_rows_limit) {
    const uint32_t logical = xcd_remap(blockIdx.x, gridDim.x);
    const int64_t ri = (int64_t)(logical % (uint64_t)nrows);
    const int64_t tile = (int64_t)(logical / (uint64_t)nrows);
    const RowDescD d = rows[ri];
    if (d.out_row >= out_rows_limit) return;
    const float* __restrict__ tail = in + d.tail_row * in_pitch;
    const bool has_head = d.head_row >= 0;
    const float* __restrict__ head = in + (has_head ? d.head_row : 0) * in_pitch;
    float* __restrict__ o = out + d.out_row * out_pitch;
    const int64_t off = d.offset;
    const int64_t base = tile * kMergeTile;
    if (VEC) {
#pragma unroll
        for (int j = 0; j < kMergeVecPerThread; ++j) {
            const int64_t t = base + (int64_t)j * (kThreads * 4) + (int64_t)threadIdx.x * 4;
            if (t + 4 <= nsamps) {
                float4a a = *reinterpret_cast<const float4a*>(tail + t);
                if (has_head) {
                    if (t >= off) {
                        const float4u h = *reinterpret_cast<const float4u*>(head + (t - off));
                        a.x = a.x + h.x;
                        a.y = a.y + h.y;
                        a.z = a.z + h.z;
                        a.w = a.w + h.w;
                    } else {
                        if (t + 0 >= off) a.x = a.x + head[t + 0 - off];
                        if (t + 1 >= off) a.y = a.y + head[t + 1 - off];
                        if (t + 2 >= off) a.z = a.z + head[t + 2 - off];
                        if (t + 3 >= off) a.w = a.w + head[t + 3 - off];
                    }
                }
                *reinterpret_cast<float4a*>(o + t) = a;
            } else {
                for (int e = 0; e < 4; ++e) {
                    const int64_t tt = t + e;
                    if (tt < nsamps) {
                        float v = tail[tt];
                        if (has_head && tt >= off) v = v + head[tt - off];
                        o[tt] = v;
                    }
                }
            }
        }
    } else {
        for (int j = 0; j < kMergeTile / kThreads; ++j) {
            const int64_t t = base + (int64_t)j * kThreads + threadIdx.x;
            if (t < nsamps) {
                float v = tail[t];
                if (has_head && t >= off) v = v + head[t - off];
                o[t] = v;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// init
// ---------------------------------------------------------------------------
template <int IN_KIND>  // 0 f32, 1 u8, 2 ones
__device__ __forceinline__ float load_in(const void* in, int64_t idx) {
    if (IN_KIND == 0) return static_cast<const float*>(in)[idx];
    if (IN_KIND == 1) return (float)static_cast<const uint8_t*>(in)[idx];
    return 1.0f;
}

template <int MODE, int IN_KIND, bool NORM>
__global__ void __launch_bounds__(kThreads) k_init(const void* __restrict__ in, int64_t in_pitch, ChanNorm norm,
                                                   int64_t nsamps, const int32_t* __restrict__ grid,
                                                   const int64_t* __restrict__ grid_off,
                                                   const int64_t* __restrict__ chan_row_off,
                                                   float* __restrict__ state, int64_t pitch) {
    const int64_t c = blockIdx.y;
    const int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (t >= nsamps) return;
    const int64_t g0 = grid_off[c];
    const int K = (int)(grid_off[c + 1] - g0);
    const int32_t* g = grid + g0;
    float* row = state + chan_row_off[c] * pitch;
    const int64_t xb = c * in_pitch;
    float mu = 0.f, sg = 1.f;
    bool keep = true;
    if (NORM) {
        keep = norm.keep[c] != 0;
        mu = norm.mu[c];
        sg = norm.sigma[c];
    }
    auto X = [&](int64_t i) -> float {
        float v = load_in<IN_KIND>(in, xb + i);
        if (NORM) v = keep ? (v - mu) / sg : 0.0f;
        return v;
    };
    const int64_t d0 = g[0];
    float prev = 0.0f;
    if (t >= d0) {
        if (MODE == kInitCFdmtD) {
            prev = X(t);
        } else {
            float sum = 0.0f;
            for (int64_t i = t - d0; i <= t; ++i) sum += X(i);
            prev = sum / (float)(d0 + 1);
        }
    }
    row[t] = prev;
    for (int k = 1; k < K; ++k) {
        const int64_t dc = g[k], dp = g[k - 1];
        float v = 0.0f;
        if (t >= dc) {
            if (MODE == kInitMeanD) {
                float sum = 0.0f;
                for (int64_t i = t - dc; i < t - dp; ++i) sum += X(i);
                v = (prev * ((float)dp + 1.0f) + sum) / ((float)dc + 1.0f);
            } else if (MODE == kInitRestrictD) {
                const float sum = 0.0f + X(t - dc);
                v = (prev * (((float)dp + 1.0f) + sum)) / ((float)dc + 1.0f);
            } else {
                v = (prev * (float)dc + X(t - dc)) / (float)(dc + 1);
            }
        }
        row[(int64_t)k * pitch + t] = v;
        prev = v;
    }
}

// ---------------------------------------------------------------------------
// adjoints
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kThreads) k_merge_adjoint(const float* __restrict__ g_out, int64_t out_pitch,
                                                            float* __restrict__ g_prev, int64_t prev_pitch,
                                                            const int64_t* __restrict__ csr_off,
                                                            const int64_t* __restrict__ csr_row,
                                                            const int64_t* __restrict__ csr_shift,
                                                            int64_t nsamps) {
    const int64_t p = blockIdx.y;
    const int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (t >= nsamps) return;
    float acc = 0.0f;
    for (int64_t k = csr_off[p]; k < csr_off[p + 1]; ++k) {
        const int64_t tt = t + csr_shift[k];
        if (tt < nsamps) acc += g_out[csr_row[k] * out_pitch + tt];
    }
    g_prev[p * prev_pitch + t] = acc;
}

__global__ void __launch_bounds__(kThreads) k_init_adjoint(const float* __restrict__ g_state, int64_t pitch,
                                                           float* __restrict__ g_x, int64_t x_pitch, int64_t nsamps,
                                                           const int32_t* __restrict__ grid,
                                                           const int64_t* __restrict__ grid_off,
                                                           const int64_t* __restrict__ chan_row_off) {
    const int64_t c = blockIdx.y;
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= nsamps) return;
    const int64_t g0 = grid_off[c];
    const int K = (int)(grid_off[c + 1] - g0);
    const float* base = g_state + chan_row_off[c] * pitch;
    float acc = 0.0f;
    for (int k = 0; k < K; ++k) {
        const int64_t d = grid[g0 + k];
        const int64_t lo = i > d ? i : d;
        const int64_t hi = (i + d < nsamps - 1) ? i + d : nsamps - 1;
        float s = 0.0f;
        for (int64_t t = lo; t <= hi; ++t) s += base[(int64_t)k * pitch + t];
        acc += s / (float)(d + 1);
    }
    g_x[c * x_pitch + i] = acc;
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
void launch_fdmt_merge(const float* in, int64_t in_pitch, float* out, int64_t out_pitch, const RowDescD* d_rows,
                       int64_t nrows, int64_t nsamps, int64_t out_rows_limit, hipStream_t st) {
    if (nrows <= 0 || nsamps <= 0) return;
    const int64_t tiles = ceil_div(nsamps, kMergeTile);
    const int64_t nb = nrows * tiles;
    if (nb > 0x7fffffffLL) throw Error(SPS_ERR_INVALID_ARGUMENT, "merge grid too large");
    const bool vec = (in_pitch % 4 == 0) && (out_pitch % 4 == 0) &&
                     (reinterpret_cast<uintptr_t>(in) % 16 == 0) && (reinterpret_cast<uintptr_t>(out) % 16 == 0);
    if (vec)
        hipLaunchKernelGGL((k_merge<true>), dim3((uint32_t)nb), dim3(kThreads), 0, st, in, in_pitch, out, out_pitch,
                           d_rows, nrows, nsamps, out_rows_limit);
    else
        hipLaunchKernelGGL((k_merge<false>), dim3((uint32_t)nb), dim3(kThreads), 0, st, in, in_pitch, out, out_pitch,
                           d_rows, nrows, nsamps, out_rows_limit);
    SPS_HIP(hipGetLastError());
}

template <int MODE, int IN_KIND>
static void init_dispatch_norm(const void* in, int64_t in_pitch, const ChanNorm* norm, int64_t nchans,
                               int64_t nsamps, const int32_t* g, const int64_t* go, const int64_t* cro, float* state,
                               int64_t pitch, hipStream_t st) {
    dim3 grid((uint32_t)ceil_div(nsamps, kThreads), (uint32_t)nchans);
    ChanNorm n{};
    if (norm) {
        n = *norm;
        hipLaunchKernelGGL((k_init<MODE, IN_KIND, true>), grid, dim3(kThreads), 0, st, in, in_pitch, n, nsamps, g, go,
                           cro, state, pitch);
    } else {
        hipLaunchKernelGGL((k_init<MODE, IN_KIND, false>), grid, dim3(kThreads), 0, st, in, in_pitch, n, nsamps, g, go,
                           cro, state, pitch);
    }
}

template <int MODE>
static void init_dispatch_kind(const void* in, int in_u8, int ones, int64_t in_pitch, const ChanNorm* norm,
                               int64_t nchans, int64_t nsamps, const int32_t* g, const int64_t* go,
                               const int64_t* cro, float* state, int64_t pitch, hipStream_t st) {
    if (ones)
        init_dispatch_norm<MODE, 2>(in, in_pitch, norm, nchans, nsamps, g, go, cro, state, pitch, st);
    else if (in_u8)
        init_dispatch_norm<MODE, 1>(in, in_pitch, norm, nchans, nsamps, g, go, cro, state, pitch, st);
    else
        init_dispatch_norm<MODE, 0>(in, in_pitch, norm, nchans, nsamps, g, go, cro, state, pitch, st);
}

void launch_fdmt_init(int mode, const void* in, int in_u8, int ones, int64_t in_pitch, const ChanNorm* norm,
                      int64_t nchans, int64_t nsamps, const int32_t* d_grid, const int64_t* d_grid_off,
                      const int64_t* d_chan_row_off, float* state, int64_t pitch, hipStream_t st) {
    if (nchans <= 0 || nsamps <= 0) return;
    if (nchans > 65535) throw Error(SPS_ERR_INVALID_ARGUMENT, "more than 65535 channels");
    switch (mode) {
        case kInitMeanD:
            init_dispatch_kind<kInitMeanD>(in, in_u8, ones, in_pitch, norm, nchans, nsamps, d_grid, d_grid_off,
                                           d_chan_row_off, state, pitch, st);
            break;
        case kInitRestrictD:
            init_dispatch_kind<kInitRestrictD>(in, in_u8, ones, in_pitch, norm, nchans, nsamps, d_grid, d_grid_off,
                                               d_chan_row_off, state, pitch, st);
            break;
        case kInitCFdmtD:
            init_dispatch_kind<kInitCFdmtD>(in, in_u8, ones, in_pitch, norm, nchans, nsamps, d_grid, d_grid_off,
                                            d_chan_row_off, state, pitch, st);
            break;
        default:
            throw Error(SPS_ERR_INVALID_ARGUMENT, "unknown init mode");
    }
    SPS_HIP(hipGetLastError());
}

void launch_fdmt_merge_adjoint(const float* g_out, int64_t out_pitch, float* g_prev, int64_t prev_pitch,
                               const int64_t* d_csr_off, const int64_t* d_csr_row, const int64_t* d_csr_shift,
                               int64_t prev_rows, int64_t nsamps, hipStream_t st) {
    if (prev_rows <= 0 || nsamps <= 0) return;
    if (prev_rows > 65535) {
        // fold rows over several launches to respect gridDim.y
        for (int64_t r0 = 0; r0 < prev_rows; r0 += 65535) {
            const int64_t nr = (prev_rows - r0 < 65535) ? prev_rows - r0 : 65535;
            dim3 grid((uint32_t)ceil_div(nsamps, kThreads), (uint32_t)nr);
            hipLaunchKernelGGL(k_merge_adjoint, grid, dim3(kThreads), 0, st, g_out, out_pitch, g_prev + r0 * prev_pitch,
                               prev_pitch, d_csr_off + r0, d_csr_row, d_csr_shift, nsamps);
        }
    } else {
        dim3 grid((uint32_t)ceil_div(nsamps, kThreads), (uint32_t)prev_rows);
        hipLaunchKernelGGL(k_merge_adjoint, grid, dim3(kThreads), 0, st, g_out, out_pitch, g_prev, prev_pitch,
                           d_csr_off, d_csr_row, d_csr_shift, nsamps);
    }
    SPS_HIP(hipGetLastError());
}

void launch_fdmt_init_adjoint(const float* g_state, int64_t pitch, float* g_x, int64_t x_pitch, int64_t nchans,
                              int64_t nsamps, const int32_t* d_grid, const int64_t* d_grid_off,
                              const int64_t* d_chan_row_off, hipStream_t st) {
    if (nchans <= 0 || nsamps <= 0) return;
    if (nchans > 65535) throw Error(SPS_ERR_INVALID_ARGUMENT, "more than 65535 channels");
    dim3 grid((uint32_t)ceil_div(nsamps, kThreads), (uint32_t)nchans);
    hipLaunchKernelGGL(k_init_adjoint, grid, dim3(kThreads), 0, st, g_state, pitch, g_x, x_pitch, nsamps, d_grid,
                       d_grid_off, d_chan_row_off);
    SPS_HIP(hipGetLastError());
}

}  // namespace sps
