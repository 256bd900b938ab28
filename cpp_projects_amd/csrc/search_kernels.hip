// search_kernels.hip -- statistics/cleaning, S/N normalisation, boxcar
// candidate gathering and the single-maximum detector for gfx950.
//
// Statistics use FP64 accumulators with a fixed lane partition and a fixed
// reduction tree, so results do not depend on launch shape or timing:
//   row sums   : element t goes to lane (t / 16) % 256, lanes accumulate in
//                increasing t, then pairwise tree over the 256 lanes;
//   cross-row  : row r goes to lane r % 256, same tree.
// For integer (u8) input every partial sum is an exact integer in FP64.
#include "common.hpp"
#include "detect.hpp"
#include "kernels.hpp"

namespace sps {

namespace {

constexpr int kLanes = 256;
constexpr int kChunk = 16;

__device__ __forceinline__ void tree_reduce2(double* a, double* b) {
    for (int s = kLanes / 2; s > 0; s >>= 1) {
        __syncthreads();
        if ((int)threadIdx.x < s) {
            a[threadIdx.x] = a[threadIdx.x] + a[threadIdx.x + s];
            b[threadIdx.x] = b[threadIdx.x] + b[threadIdx.x + s];
        }
    }
    __syncthreads();
}

template <bool U8>
__global__ void __launch_bounds__(kLanes) k_row_stats(const void* __restrict__ in, int64_t cols,
                                                      double* __restrict__ row_mean, double* __restrict__ row_var) {
    __shared__ double s1[kLanes], s2[kLanes];
    const int64_t r = blockIdx.x;
    double a = 0.0, b = 0.0;
    const int64_t nchunks = ceil_div(cols, kChunk);
    int64_t ch0 = threadIdx.x;
    if (!U8) {
        // four 64-byte chunks (16 floats each, four float4 loads) in flight
        // per lane, the lane's elements still added in increasing t
        const float* x = static_cast<const float*>(in) + r * cols;
        if ((reinterpret_cast<uintptr_t>(x) & 15) == 0) {
            const int64_t nfull = cols / kChunk;
            for (; ch0 + 3 * kLanes < nfull; ch0 += 4 * kLanes) {
                float4 v[4][kChunk / 4];
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int q = 0; q < kChunk / 4; ++q)
                        v[u][q] = *reinterpret_cast<const float4*>(x + (ch0 + u * kLanes) * kChunk + 4 * q);
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int q = 0; q < kChunk / 4; ++q) {
                        const float e4[4] = {v[u][q].x, v[u][q].y, v[u][q].z, v[u][q].w};
#pragma unroll
                        for (int e = 0; e < 4; ++e) {
                            const double d = (double)e4[e];
                            a += d;
                            b += d * d;
                        }
                    }
            }
        }
    }
    if (U8) {
        // four 16-byte loads in flight per lane: one per iteration left the
        // kernel waiting on load latency.  u8 sums and sums of squares are
        // integers below 2^53 for rows under 1.3e11 samples, exact in FP64,
        // so neither the grouping nor integer arithmetic changes them: the
        // 64 samples of an iteration go through 4-byte integer dot products
        // (sums <= 16320 and sums of squares <= 4161600 fit 32 bits) into
        // 64-bit lane totals, converted once
        const uint8_t* x = static_cast<const uint8_t*>(in) + r * cols;
        if ((reinterpret_cast<uintptr_t>(x) & 15) == 0) {
            const int64_t nfull = cols / kChunk;
            uint64_t ai = 0, bi = 0;
            for (; ch0 + 3 * kLanes < nfull; ch0 += 4 * kLanes) {
                uint4 v[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) v[u] = *reinterpret_cast<const uint4*>(x + (ch0 + u * kLanes) * kChunk);
                uint32_t s = 0u, q = 0u;
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const uint32_t w[4] = {v[u].x, v[u].y, v[u].z, v[u].w};
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        s = __builtin_amdgcn_udot4(w[e], 0x01010101u, s, false);
                        q = __builtin_amdgcn_udot4(w[e], w[e], q, false);
                    }
                }
                ai += s;
                bi += q;
            }
            a = (double)ai;
            b = (double)bi;
        }
    }
    for (int64_t ch = ch0; ch < nchunks; ch += kLanes) {
        const int64_t t0 = ch * kChunk;
        if (U8) {
            const uint8_t* x = static_cast<const uint8_t*>(in) + r * cols;
            if (t0 + kChunk <= cols && ((reinterpret_cast<uintptr_t>(x + t0) & 15) == 0)) {
                const uint4 v = *reinterpret_cast<const uint4*>(x + t0);
                const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int q = 0; q < 4; ++q)
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const double d = (double)((w[q] >> (8 * e)) & 0xffu);
                        a += d;
                        b += d * d;
                    }
            } else {
                for (int64_t t = t0; t < t0 + kChunk && t < cols; ++t) {
                    const double d = (double)x[t];
                    a += d;
                    b += d * d;
                }
            }
        } else {
            const float* x = static_cast<const float*>(in) + r * cols;
            if (t0 + kChunk <= cols && ((reinterpret_cast<uintptr_t>(x + t0) & 15) == 0)) {
                // a whole aligned chunk: four float4 loads, added in order
                float4 v[kChunk / 4];
#pragma unroll
                for (int q = 0; q < kChunk / 4; ++q) v[q] = *reinterpret_cast<const float4*>(x + t0 + 4 * q);
#pragma unroll
                for (int q = 0; q < kChunk / 4; ++q) {
                    const float e4[4] = {v[q].x, v[q].y, v[q].z, v[q].w};
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const double d = (double)e4[e];
                        a += d;
                        b += d * d;
                    }
                }
            } else {
                for (int64_t t = t0; t < t0 + kChunk && t < cols; ++t) {
                    const double d = (double)x[t];
                    a += d;
                    b += d * d;
                }
            }
        }
    }
    s1[threadIdx.x] = a;
    s2[threadIdx.x] = b;
    tree_reduce2(s1, s2);
    if (threadIdx.x == 0) {
        const double m = s1[0] / (double)cols;
        row_mean[r] = m;
        row_var[r] = s2[0] / (double)cols - m * m;
    }
}

__global__ void __launch_bounds__(kLanes) k_stats_finalize(const double* __restrict__ rm, const double* __restrict__ rv,
                                                           int64_t rows, int mode, double* __restrict__ stats,
                                                           float* __restrict__ mu, float* __restrict__ sigma,
                                                           uint8_t* __restrict__ keep, float* __restrict__ buff) {
    __shared__ double a[kLanes], b[kLanes], c[kLanes], d[kLanes];
    __shared__ double res[5];
    double sa = 0.0, sb = 0.0, sc = 0.0, sd = 0.0;
    for (int64_t r = threadIdx.x; r < rows; r += kLanes) {
        sa = sa + rm[r];
        sb = sb + (rv[r] + rm[r] * rm[r]);
        sc = sc + rv[r];
        sd = sd + rv[r] * rv[r];
    }
    a[threadIdx.x] = sa;
    b[threadIdx.x] = sb;
    c[threadIdx.x] = sc;
    d[threadIdx.x] = sd;
    tree_reduce2(a, b);
    tree_reduce2(c, d);
    if (threadIdx.x == 0) {
        const double mean = a[0] / (double)rows;
        const double sdv = sqrt(b[0] / (double)rows - mean * mean);
        const double vm = c[0] / (double)rows;
        const double vs = sqrt(d[0] / (double)rows - vm * vm);
        res[0] = mean;
        res[1] = sdv;
        res[2] = vm;
        res[3] = vs;
    }
    __syncthreads();
    uint32_t nk = 0;
    for (int64_t r = threadIdx.x; r < rows; r += kLanes) {
        bool drop = fabs(rv[r] - res[2]) > 4.0 * res[3];
        float m, s;
        if (mode == 2) {
            if (!(rv[r] > 0.0)) drop = true;
            m = (float)rm[r];
            s = (float)sqrt(rv[r]);
        } else {
            m = (float)res[0];
            s = (float)res[1];
        }
        keep[r] = drop ? 0 : 1;
        mu[r] = m;
        sigma[r] = s;
        nk += drop ? 0u : 1u;
        if (buff) {
            buff[r] = (float)rm[r];
            buff[rows + r] = (float)rv[r];
        }
    }
    __shared__ uint32_t cnt;
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    atomicAdd(&cnt, nk);
    __syncthreads();
    if (threadIdx.x == 0) {
        stats[0] = res[0];
        stats[1] = res[1];
        stats[2] = res[2];
        stats[3] = res[3];
        stats[4] = (double)cnt;
        stats[5] = 0.0;
        if (buff) {
            buff[2 * rows + 0] = (float)res[0];
            buff[2 * rows + 1] = (float)res[1];
            buff[2 * rows + 2] = (float)res[2];
            buff[2 * rows + 3] = (float)res[3];
        }
    }
}

__global__ void __launch_bounds__(kThreads) k_apply_norm(float* __restrict__ x, float* __restrict__ mask, int64_t cols,
                                                         const float* __restrict__ mu, const float* __restrict__ sigma,
                                                         const uint8_t* __restrict__ keep) {
    const int64_t r = blockIdx.y;
    const bool k = keep[r] != 0;
    const float m = mu[r], s = sigma[r];
    for (int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x; t < cols; t += (int64_t)gridDim.x * kThreads) {
        const int64_t i = r * cols + t;
        x[i] = k ? (x[i] - m) / s : 0.0f;
        if (mask) mask[i] = k ? 1.0f : 0.0f;
    }
}

__global__ void __launch_bounds__(kThreads) k_fdmt_normalization(const float* __restrict__ a,
                                                                 const float* __restrict__ n, int64_t len,
                                                                 float* __restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < len; i += (int64_t)gridDim.x * kThreads)
        out[i] = a[i] / sqrtf((float)((double)n[i] + 1.0E-8));
}

// fdmt_normalization with a causal norm image kept for its first norm_cols
// columns: out[r][t] = a[r][t] / sqrtf(norm[r][min(t, norm_cols - 1)] + 1e-8)
__global__ void __launch_bounds__(kThreads) k_fdmt_normalization_rows(const float* __restrict__ a, int64_t cols,
                                                                      const float* __restrict__ n, int64_t norm_cols,
                                                                      float* __restrict__ out) {
    const int64_t r = blockIdx.y;
    const float* nr = n + r * norm_cols;
    const float tail = sqrtf((float)((double)nr[norm_cols - 1] + 1.0E-8));
    for (int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x; t < cols; t += (int64_t)gridDim.x * kThreads) {
        const float d = t < norm_cols ? sqrtf((float)((double)nr[t] + 1.0E-8)) : tail;
        out[r * cols + t] = a[r * cols + t] / d;
    }
}

__global__ void __launch_bounds__(kThreads) k_snr_normalise(float* __restrict__ img, int64_t cols, int64_t pitch,
                                                            const float* __restrict__ norm, int64_t norm_len,
                                                            const double* __restrict__ stats) {
    const int64_t r = blockIdx.y;
    const float nk = (float)stats[4];
    const float dk = sqrtf((float)((double)nk + 1.0E-8));
    for (int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x; t < cols; t += (int64_t)gridDim.x * kThreads) {
        float* p = img + r * pitch + t;
        if (t < norm_len)
            *p = *p / sqrtf((float)((double)norm[r * norm_len + t] + 1.0E-8));
        else
            *p = *p / dk;
    }
}

// ---------------------------------------------------------------------------
// boxcar: exact left-to-right window sums (Clusterization.cu:419-462,591-671)
// ---------------------------------------------------------------------------
// Candidates are staged per block in LDS and reserved in global memory with
// one atomic per block; a full stage spills through direct atomics.
constexpr int kStage = 512;

struct CandStage {
    CandD c[kStage];
    uint32_t n;
    unsigned long long base;
};

__device__ __forceinline__ void stage_emit(CandStage& st, CandD c, CandD* __restrict__ out, int64_t cap,
                                           unsigned long long* __restrict__ count) {
    const uint32_t k = atomicAdd(&st.n, 1u);
    if (k < (uint32_t)kStage) {
        st.c[k] = c;
    } else {
        const unsigned long long g = atomicAdd(count, 1ull);
        if ((int64_t)g < cap) out[g] = c;
    }
}

__device__ __forceinline__ void stage_flush(CandStage& st, CandD* __restrict__ out, int64_t cap,
                                            unsigned long long* __restrict__ count) {
    __syncthreads();
    const uint32_t n = min(st.n, (uint32_t)kStage);
    if (threadIdx.x == 0) st.base = n ? atomicAdd(count, (unsigned long long)n) : 0ull;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += kThreads) {
        const unsigned long long g = st.base + i;
        if ((int64_t)g < cap) out[g] = st.c[i];
    }
}

// All widths 1..wmax: the running left-to-right sum is exact for every width.
constexpr int kBoxCells = 1024;  // cells per block (4 per thread)

// Time tile of a (tile, row) grid, rotated by the row.  The hardware deals
// consecutive block ids round-robin to the 8 XCDs, so with a multiple of 8
// tiles per row, tile t of every row would run on XCD t % 8; candidate-dense
// blocks follow tracks across rows (a pulse's arrival drifts slowly with DM)
// and would pile onto one XCD while the others idle.  Rotating by the row
// (a bijection within each row) spreads such a track over all XCDs.
__device__ __forceinline__ int64_t rotated_tile() {
    return ((int64_t)blockIdx.x + (int64_t)blockIdx.y) % (int64_t)gridDim.x;
}

__global__ void __launch_bounds__(kThreads) k_boxcar_all(const float* __restrict__ img, int64_t cols, int64_t pitch,
                                                         const float* __restrict__ d_thr, float h_thr, int wmax,
                                                         CandD* __restrict__ out, int64_t cap,
                                                         unsigned long long* __restrict__ count) {
    extern __shared__ __attribute__((aligned(16))) float tile[];
    __shared__ CandStage st;
    if (threadIdx.x == 0) st.n = 0;
    const int64_t r = blockIdx.y;
    const int64_t T = rotated_tile() * kBoxCells;
    const int64_t span = min((int64_t)kBoxCells + wmax - 1, cols - T);
    const float* row = img + r * pitch;
    for (int64_t i = threadIdx.x; i < span; i += kThreads) tile[i] = row[T + i];
    __syncthreads();
    const float thr = d_thr ? *d_thr : h_thr;
    for (int j = 0; j < kBoxCells / kThreads; ++j) {
        const int64_t lt = (int64_t)j * kThreads + threadIdx.x;
        const int64_t t = T + lt;
        if (t >= cols) break;
        const int64_t wlim = min((int64_t)wmax, cols - t);
        float s = 0.0f;
        for (int64_t iw = 0; iw < wlim; ++iw) {
            s += tile[lt + iw];
            const float snr = s / sqrtf((float)(iw + 1));
            if (snr > thr) stage_emit(st, CandD{(int32_t)t, (int32_t)r, (int32_t)(iw + 1), snr}, out, cap, count);
        }
    }
    stage_flush(st, out, cap, count);
}

// Power-of-two widths 1..2^q.  Exactness with a fast filter:
//  * dyadic pairwise sums P_k(t) = P_{k-1}(t) + P_{k-1}(t + 2^{k-1}) in LDS;
//  * the reference value is the left-to-right fp32 sum s; both are within
//    gamma_{w-1}*sum|x| resp. gamma_k*sum|x| of the exact sum, so
//    |P - s| <= (gamma_{w-1} + gamma_k) * w * M with M = max|x| over the
//    block's samples (gamma_n = n u / (1 - n u), u = 2^-24);
//  * fl(s / sqrtf(w)) > thr needs s > thr*sqrtf(w)/(1+u), so a cell whose
//    P is below that bound minus the error term cannot pass and is skipped;
//  * every other cell is recomputed left to right (one running sum up to its
//    largest flagged width) and decided exactly as the reference does.
constexpr int kPowCells = 2048;            // cells per block (8 per thread)
constexpr int kPowPer = kPowCells / kThreads;
constexpr int kMaxPowQ = 13;

struct BoxErr {
    double e[kMaxPowQ + 1];
};

// wave-wide max from the ROCm device library (DPP within rows, then across
// rows): a short VALU chain instead of six LDS-routed shuffles
extern "C" __device__ float __ockl_wfred_max_f32(float);
extern "C" __device__ int __ockl_wfred_max_i32(int);

// threadIdx.x through an empty asm statement: the compiler cannot hoist
// what is computed from it out of a loop.  The boxcar tile body uses it, so
// that a block looping over tiles recomputes its per-thread LDS addresses
// per tile instead of holding them in registers across the loop.
__device__ __forceinline__ unsigned opaque_tid() {
    unsigned t = threadIdx.x;
    asm volatile("" : "+v"(t));
    return t;
}

__host__ __device__ constexpr int round_up4(int v) { return (v + 3) & ~3; }

// Largest float not above the filter limit of width 2^k:
// thr*sqrtf(w)*(1 - 2^-22) - e_k*M, e_k = (gamma_{w-1}*w + E_k)*(1 + 2^-20)
// (computed on the host): gamma_{w-1}*w*M bounds the error of the reference's
// left-to-right sum, E_k*M that of the value the filter compares (gamma_k*w
// for the pairwise level sums, gamma_{k+6}*(w + 16) for the segment bound of
// widths >= 16).  For a float P, P > this float <=> P > the limit;
// NaN thresholds or NaN/infinite samples flag everything.
__device__ __forceinline__ float level_limit(float thr, int k, double e_k, float M) {
    if (!(thr > 0.0f)) return -INFINITY;
    const double lim = (double)thr * (double)sqrtf((float)(1 << k)) * (1.0 - 0x1p-22) - e_k * (double)M;
    return __double2float_rd(lim);
}

constexpr int kPowStage = 128;  // candidate stage entries (uint2) per wave

// Copies samples [0, span) of a window into LDS; returns the thread's max
// |x| (+inf once it saw a NaN, which disables the filter).  vec: 16-byte
// loads (the window start is 16-byte aligned), the last partial vector scalar.
__device__ __forceinline__ float pow2_load_window(float* x, const float* src, int span, bool vec) {
    const unsigned tid = opaque_tid();
    float m = 0.0f;
    if (vec) {
        const int n4 = span >> 2;
        const float4* src4 = reinterpret_cast<const float4*>(src);
        for (int i = tid; i < n4; i += kThreads) {
            const float4 v = src4[i];
            reinterpret_cast<float4*>(x)[i] = v;
            m = fmaxf(m, fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fmaxf(fabsf(v.z), fabsf(v.w))));
            if (v.x != v.x || v.y != v.y || v.z != v.z || v.w != v.w) m = __int_as_float(0x7f800000);
        }
        for (int i = 4 * n4 + tid; i < span; i += kThreads) {
            const float v = src[i];
            x[i] = v;
            m = fmaxf(m, fabsf(v));
            if (v != v) m = __int_as_float(0x7f800000);
        }
    } else {
        for (int i = tid; i < span; i += kThreads) {
            const float v = src[i];
            x[i] = v;
            m = fmaxf(m, fabsf(v));
            if (v != v) m = __int_as_float(0x7f800000);
        }
    }
    return m;
}

// Width 1 from the samples, then two dyadic levels per round: from level k0
// (cur), level k0+1 at i is cur[i] + cur[i+h] and level k0+2 is that plus
// cur[i+2h] + cur[i+3h] (h = 2^k0), the pairwise sums of the separate
// levels.  Only level k0+2 is stored; each round reads one buffer and writes
// the other.  All reads of a round are issued before any is used (no
// branches between them): indexes stay below span_max (i + 3h <= 2047 + 3W/4
// when two levels are formed, i + h <= 2047 + W/2 otherwise) and entries at
// or past a level's valid span are computed but never flagged or consumed.
// flags[j] = f with f <- 2f + (level k flagged) for k = 0..q.  EDGE: the
// window is cut by the row end, so a level is valid at i only while
// i + 2^k <= span (span <= cols - T: a window inside the span is inside the
// row); otherwise span = 2048 + W - 1 and every level is valid at every own
// cell.
template <bool EDGE>
__device__ __forceinline__ void pow2_flag_rounds(float* x, float* la, int q, int span, const float* lim,
                                                 uint32_t (&flags)[kPowPer]) {
    const unsigned tid = opaque_tid();
    {
        const float lf = lim[0];
        float v[kPowPer];
#pragma unroll
        for (int j = 0; j < kPowPer; ++j) v[j] = x[j * kThreads + tid];
#pragma unroll
        for (int j = 0; j < kPowPer; ++j) {
            const int lt = j * kThreads + tid;
            flags[j] = ((!EDGE || lt < span) && !(v[j] <= lf)) ? 1u : 0u;
        }
    }
    const float* cur = x;
    for (int k0 = 0; k0 < q; k0 += 2) {
        const int k1 = k0 + 1, k2 = k0 + 2;
        const int h = 1 << k0;
        const float l1 = lim[k1];
        const bool two = k2 <= q;
        const float l2 = two ? lim[k2] : 0.0f;
        // ping-pong: round 0 reads the samples and writes la, round 1 reads
        // la and writes over the samples (read by round 0 only), and so on
        float* dst = ((k0 >> 1) & 1) ? x : la;
        float va[kPowPer], vb[kPowPer], vc[kPowPer], vd[kPowPer];
        if (two) {
#pragma unroll
            for (int j = 0; j < kPowPer; ++j) {
                const int i = j * kThreads + tid;
                va[j] = cur[i];
                vb[j] = cur[i + h];
                vc[j] = cur[i + 2 * h];
                vd[j] = cur[i + 3 * h];
            }
#pragma unroll
            for (int j = 0; j < kPowPer; ++j) {  // the thread's own cells
                const int i = j * kThreads + tid;
                const float p1 = va[j] + vb[j];
                const float p2 = p1 + (vc[j] + vd[j]);
                dst[i] = p2;
                const bool c1 = (!EDGE || i + 2 * h <= span) && !(p1 <= l1);
                const bool c2 = (!EDGE || i + 4 * h <= span) && !(p2 <= l2);
                flags[j] = flags[j] + flags[j] + (uint32_t)c1;
                flags[j] = flags[j] + flags[j] + (uint32_t)c2;
            }
        } else {
#pragma unroll
            for (int j = 0; j < kPowPer; ++j) {
                const int i = j * kThreads + tid;
                va[j] = cur[i];
                vb[j] = cur[i + h];
            }
#pragma unroll
            for (int j = 0; j < kPowPer; ++j) {
                const int i = j * kThreads + tid;
                const float p1 = va[j] + vb[j];
                const bool c1 = (!EDGE || i + 2 * h <= span) && !(p1 <= l1);
                flags[j] = flags[j] + flags[j] + (uint32_t)c1;
            }
        }
        if (two) {  // the halo: stored for the next rounds only
            for (int i = kPowCells + tid; i + 4 * h <= span; i += kThreads)
                dst[i] = (cur[i] + cur[i + h]) + (cur[i + 2 * h] + cur[i + 3 * h]);
        }
        __syncthreads();
        cur = dst;
    }
}

__device__ __forceinline__ float max4(const float4& v) { return fmaxf(fmaxf(v.x, v.y), fmaxf(v.z, v.w)); }

// The same rounds in the chunk layout, for interior windows (span = 2048 +
// W - 1, so every level is valid at every own cell) whose samples are finite
// with W * M < 2^125 (no level sum is NaN or overflows).  Thread t owns chunks
// t and t + 256 of four consecutive cells; every LDS access is 16 bytes.  A
// level is flagged for a chunk when the largest of its four sums exceeds the
// limit, and the chunk's four cells take those flags (a superset of the
// per-cell flags: the recompute decides each cell exactly).  The level sums
// are those of pow2_flag_rounds, in the same order.  Reads stay below
// span_max: chunk 511 + 3h/4 (cell 2047 + 3h <= 2047 + 3W/4) when two levels
// are formed, 511 + h/4 otherwise, and chunk 512 (cells 2048..2051 <
// span_max for W >= 2) in the first round.
// Q is a template parameter so that the rounds unroll: every round's
// branches (first round, one or two levels) resolve at compile time and no
// values are carried across rounds.
template <int Q>
__device__ __forceinline__ void pow2_flag_chunks(float* x, float* la, const float* lim, uint32_t (&cf)[2]) {
    const unsigned tid = opaque_tid();
    constexpr int q = Q;
    constexpr int span = kPowCells + (1 << q) - 1;
    {
        const float lf = lim[0];
        const float4* x4 = reinterpret_cast<const float4*>(x);
#pragma unroll
        for (int s = 0; s < 2; ++s) cf[s] = !(max4(x4[s * kThreads + tid]) <= lf) ? 1u : 0u;
    }
    const float* cur = x;
#pragma unroll
    for (int k0 = 0; k0 < q; k0 += 2) {
        const int k1 = k0 + 1, k2 = k0 + 2;
        const int h = 1 << k0;
        const float l1 = lim[k1];
        const bool two = k2 <= q;
        const float l2 = two ? lim[k2] : 0.0f;
        float* dst = ((k0 >> 1) & 1) ? x : la;
        const float4* c4 = reinterpret_cast<const float4*>(cur);
        float4* d4 = reinterpret_cast<float4*>(dst);
        float4 p1[2], p2[2];
        if (k0 == 0) {
            // h = 1: cells 4c .. 4c+7 cover both levels of chunk c
            float4 w0[2], w1[2];
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                const int c = s * kThreads + tid;
                w0[s] = c4[c];
                w1[s] = c4[c + 1];
            }
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                const float4 a = w0[s], b = w1[s];
                p1[s] = make_float4(a.x + a.y, a.y + a.z, a.z + a.w, a.w + b.x);
                p2[s] = make_float4(p1[s].x + (a.z + a.w), p1[s].y + (a.w + b.x), p1[s].z + (b.x + b.y),
                                    p1[s].w + (b.y + b.z));
            }
        } else {
            // h a multiple of 4: the four terms are whole chunks h/4 apart
            const int hq = h >> 2;
            float4 A[2], B[2], C[2], D[2];
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                const int c = s * kThreads + tid;
                A[s] = c4[c];
                B[s] = c4[c + hq];
                if (two) {
                    C[s] = c4[c + 2 * hq];
                    D[s] = c4[c + 3 * hq];
                }
            }
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                p1[s] = make_float4(A[s].x + B[s].x, A[s].y + B[s].y, A[s].z + B[s].z, A[s].w + B[s].w);
                if (two)
                    p2[s] = make_float4(p1[s].x + (C[s].x + D[s].x), p1[s].y + (C[s].y + D[s].y),
                                        p1[s].z + (C[s].z + D[s].z), p1[s].w + (C[s].w + D[s].w));
            }
        }
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            cf[s] = cf[s] + cf[s] + (!(max4(p1[s]) <= l1) ? 1u : 0u);
            if (two) {
                d4[s * kThreads + tid] = p2[s];
                cf[s] = cf[s] + cf[s] + (!(max4(p2[s]) <= l2) ? 1u : 0u);
            }
        }
        if (two) {  // the halo: stored for the next rounds only
            for (int i = kPowCells + tid; i + 4 * h <= span; i += kThreads)
                dst[i] = (cur[i] + cur[i + h]) + (cur[i + 2 * h] + cur[i + 3 * h]);
        }
        __syncthreads();
        cur = dst;
    }
}

// Runs pow2_flag_chunks<q> for a runtime q in [Q, QMAX].
template <int Q, int QMAX>
__device__ __forceinline__ void pow2_flag_chunks_at(int q, float* x, float* la, const float* lim, uint32_t (&cf)[2]) {
    if (q == Q) {
        pow2_flag_chunks<Q>(x, la, lim, cf);
    } else {
        if constexpr (Q < QMAX) pow2_flag_chunks_at<Q + 1, QMAX>(q, x, la, lim, cf);
    }
}

constexpr int kPowPfMaxQ = 9;  // widest window of the segment filter and the multi-tile kernel

// max of four values (the chunk filter's flag test), as in pow2_flag_chunks
__device__ __forceinline__ uint32_t chunk_flag(float a, float b, float c, float d, float l) {
    return !(fmaxf(fmaxf(a, b), fmaxf(c, d)) <= l) ? 1u : 0u;
}

// Filter limit of width 2^k from the block's coefficients lc[k] = {A_k, E_k}
// (A_k <= thr*sqrtf(w)*(1 - 2^-22) and E_k >= e_k, both rounded on the safe
// side; see level_limit): t = RN(A_k - E_k*M) is within half an ulp of the
// exact value, so t - |t|*2^-22 - 2^-100 (each step rounded) lies below
// A_k - E_k*M <= the limit of level_limit.  A lower limit only flags more
// cells; the recompute decides each exactly.
__device__ __forceinline__ float seg_limit(float2 lc, float M) {
    const float t = fmaf(-lc.y, M, lc.x);
    return fmaf(-fabsf(t), 0x1p-22f, t) - 0x1p-100f;
}

// The block's limit coefficients for widths 2^0 .. 2^q (thread k computes
// level k's; the caller makes them visible with a barrier).
__device__ __forceinline__ float2 seg_coef(float thr, int k, double e_k) {
    if (!(thr > 0.0f)) return make_float2(-INFINITY, 0.0f);
    const double a = (double)thr * (double)sqrtf((float)(1 << k)) * (1.0 - 0x1p-22);
    return make_float2(__double2float_rd(a), __double2float_ru(e_k));
}

// The divisor of width 2^p, sqrtf(2^p) (IEEE), and its reciprocal rounded to
// nearest (exactly 2^(-p/2) for even p)
__device__ __forceinline__ float2 box_divisor(int p) {
    const float d = sqrtf((float)(1 << p));
    return make_float2(d, 1.0f / d);
}

// |v| as an ordered integer: finite magnitudes < 0x7f800000 (inf) < NaNs
__device__ __forceinline__ uint32_t abs_bits(float v) { return __float_as_uint(v) & 0x7fffffffu; }

// The flag phase of an interior window (span = 2048 + W - 1, every level
// valid at every own cell) with thread t owning segment t = cells 8t ..
// 8t+7, i.e. chunks 2t (cf[0]) and 2t + 1 (cf[1]).  It also takes the
// window's max |x| M from the cells it loads (the cells of the DMA'd buffer
// [0, 2048 + W) for W >= 16, of [0, 2064) below; cells past the window only
// raise M) with one barrier, then the limits from M.  Returns false, flags
// unset, when the window needs the careful per-cell path: NaN or infinite
// samples, or W * M >= 2^125 (a level sum could overflow); M is returned
// either way.
//  * Widths 1 .. 8: the per-cell pairwise level sums of pow2_flag_rounds
//    (P_2w(i) = P_w(i) + P_w(i + w)), from the thread's own cells and the
//    next 7; a chunk is flagged when its largest sum exceeds the limit.
//  * Widths w = 2^k >= 16: a window starting at cell 8g + m (m < 8) covers
//    segments g .. g + w/8 - 1 minus the first m cells of g, plus the first
//    m cells of segment g' = g + w/8.  With pre_g(m) the sum of segment g's
//    first m cells (pre_g(0) = 0) and D(g) the sum of segments g ..
//    g + w/8 - 1, it is D(g) - pre_g(m) + pre_g'(m) <= D(g) + max_m pre_g'(m)
//    - min_m pre_g(m) = U(g), so both chunks of segment g are flagged at level
//    k when U(g) exceeds the limit (the host's error term for k >= 4 covers
//    the rounding of U: segment sums of 8, D a depth-(k-3) tree of them,
//    prefixes of <= 7, two more operations).  D doubles over segment sums
//    (D_2n(g) = D_n(g) + D_n(g + n)), one barrier per width.
// Unlike the level rounds it keeps the samples; LDS traffic per cell is a
// few floats instead of about two per level.  aux: 5 (256 + W/8) floats;
// wbits: one word per wave.
template <int Q>
__device__ __forceinline__ bool pow2_flag_segments(const float* x, float* aux, const float2* lc, uint32_t* wbits,
                                                   uint32_t (&cf)[2], float& M) {
    static_assert(Q <= kPowPfMaxQ, "segment filter window");  // the halo segments fit one wave
    const unsigned tid = opaque_tid();
    const float4* x4 = reinterpret_cast<const float4*>(x);
    float c[16];
    {
        const float4 v0 = x4[2 * tid], v1 = x4[2 * tid + 1];
        c[0] = v0.x, c[1] = v0.y, c[2] = v0.z, c[3] = v0.w, c[4] = v1.x, c[5] = v1.y, c[6] = v1.z, c[7] = v1.w;
        if constexpr (Q >= 1) {  // cells 8t + 8 .. 8t + 15 (< 2056, inside the buffer); past the window only for
                                 // sums of levels above Q, which flag nothing
            const float4 v2 = x4[2 * tid + 2], v3 = x4[2 * tid + 3];
            c[8] = v2.x, c[9] = v2.y, c[10] = v2.z, c[11] = v2.w, c[12] = v3.x, c[13] = v3.y, c[14] = v3.z,
            c[15] = v3.w;
        }
    }
    // M over exactly the window's cells: the own segments (cells 0 .. 2047),
    // then cells 2048 .. 2046 + W from the last thread's next 8 (W <= 8) or
    // from the halo segments
    uint32_t mb = 0u;
#pragma unroll
    for (int i = 0; i < 8; ++i) mb = max(mb, abs_bits(c[i]));
    if constexpr (Q >= 1 && Q < 4) {
        if (tid == kThreads - 1) {
#pragma unroll
            for (int i = 0; i < (1 << Q) - 1; ++i) mb = max(mb, abs_bits(c[8 + i]));
        }
    }
    // the chunk maxima of each level, compared once the limits are known
    float lv[4][2];
    auto max8 = [&](const float* v, int k) {
        lv[k][0] = fmaxf(fmaxf(v[0], v[1]), fmaxf(v[2], v[3]));
        lv[k][1] = fmaxf(fmaxf(v[4], v[5]), fmaxf(v[6], v[7]));
    };
    max8(c, 0);
    if constexpr (Q >= 1) {
        float p1[14], p2[12], p3[8];
#pragma unroll
        for (int i = 0; i < 14; ++i) p1[i] = c[i] + c[i + 1];
        max8(p1, 1);
        if constexpr (Q >= 2) {
#pragma unroll
            for (int i = 0; i < 12; ++i) p2[i] = p1[i] + p1[i + 2];
            max8(p2, 2);
        }
        if constexpr (Q >= 3) {
#pragma unroll
            for (int i = 0; i < 8; ++i) p3[i] = p2[i] + p2[i + 4];
            max8(p3, 3);
        }
    }
    constexpr int NQ = Q >= 4 ? (1 << Q) / 8 : 1;  // segments of the widest window
    constexpr int G = kPowCells / 8 + NQ;
    float* S = aux;
    float* MN = aux + G;
    float* MX = aux + 2 * G;
    if constexpr (Q >= 4) {
        auto stats = [&](const float* v, int g) {
            float pre = v[0];
            float mn = fminf(0.0f, pre), mx = fmaxf(0.0f, pre);
#pragma unroll
            for (int m = 1; m < 7; ++m) {
                pre = pre + v[m];
                mn = fminf(mn, pre);
                mx = fmaxf(mx, pre);
            }
            S[g] = pre + v[7];
            MN[g] = mn;
            MX[g] = mx;
        };
        stats(c, (int)tid);
        if ((int)tid < NQ) {  // halo segments 256 .. 255 + NQ (cells < 2048 + W, inside the buffer)
            const int g = kPowCells / 8 + (int)tid;
            const float4 v0 = x4[2 * g], v1 = x4[2 * g + 1];
            const float h[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
            stats(h, g);
            const int hn = ((int)tid == NQ - 1) ? 7 : 8;  // cell 2047 + W is past the window
#pragma unroll
            for (int i = 0; i < 8; ++i)
                if (i < hn) mb = max(mb, abs_bits(h[i]));
        }
    }
    mb = (uint32_t)__ockl_wfred_max_i32((int)mb);  // every value is < 2^31
    if ((tid & 63) == 0) wbits[tid >> 6] = mb;
    __syncthreads();
    uint32_t Mb = wbits[0];
#pragma unroll
    for (int i = 1; i < kThreads / 64; ++i) Mb = max(Mb, wbits[i]);
    M = __uint_as_float(min(Mb, 0x7f800000u));  // a NaN counts as +inf
    if (!(M * (float)(1 << Q) < 0x1p125f)) return false;  // block-uniform
    // from Q = 5 on, thread k computes limit k once for the block; the D
    // levels' first barrier publishes them and every comparison comes after it
    __shared__ float limv[kPowPfMaxQ + 1];
    if constexpr (Q >= 5) {
        if ((int)tid <= Q) limv[tid] = seg_limit(lc[tid], M);
    }
    auto limit = [&](int k) -> float {
        if constexpr (Q >= 5) return limv[k];
        return seg_limit(lc[k], M);
    };
    float u[Q >= 4 ? Q - 3 : 1];
    if constexpr (Q >= 4) {
        float* DB[2] = {aux + 3 * G, aux + 4 * G};
        const float mn_own = MN[tid];
        // D_n(g) is kept for g = t and, while needed by wider levels, for
        // g = t + 256 (g <= 255 + NQ - n)
        float d0 = S[tid] + S[tid + 1];
        float d1 = ((int)tid <= NQ - 3) ? S[tid + 256] + S[tid + 257] : 0.0f;
#pragma unroll
        for (int k = 4; k <= Q; ++k) {
            const int n = 1 << (k - 3);
            if (k > 4) {
                const float* prev = DB[k & 1];
                d0 = d0 + prev[tid + n / 2];
                if ((int)tid <= NQ - n - 1) d1 = d1 + prev[tid + 256 + n / 2];
            }
            u[k - 4] = (d0 + MX[tid + n]) - mn_own;
            if (k < Q) {
                float* nxt = DB[(k + 1) & 1];
                nxt[tid] = d0;
                if ((int)tid <= NQ - n - 1) nxt[tid + 256] = d1;
                __syncthreads();
            }
        }
    }
    cf[0] = 0u;
    cf[1] = 0u;
#pragma unroll
    for (int k = 0; k <= (Q < 3 ? Q : 3); ++k) {
        const float l = limit(k);
        cf[0] = cf[0] + cf[0] + (!(lv[k][0] <= l) ? 1u : 0u);
        cf[1] = cf[1] + cf[1] + (!(lv[k][1] <= l) ? 1u : 0u);
    }
#pragma unroll
    for (int k = 4; k <= Q; ++k) {
        const uint32_t f = !(u[k - 4] <= limit(k)) ? 1u : 0u;
        cf[0] = cf[0] + cf[0] + f;
        cf[1] = cf[1] + cf[1] + f;
    }
    return true;
}

// Runs pow2_flag_segments<q> for a runtime q in [Q, QM].
template <int Q, int QM>
__device__ __forceinline__ bool pow2_flag_segments_at(int q, const float* x, float* aux, const float2* lc,
                                                      uint32_t* wbits, uint32_t (&cf)[2], float& M) {
    if (q == Q) return pow2_flag_segments<Q>(x, aux, lc, wbits, cf, M);
    if constexpr (Q < QM) return pow2_flag_segments_at<Q + 1, QM>(q, x, aux, lc, wbits, cf, M);
    return false;
}

// Recompute layout after the segment filter: one lane per flagged cell
// (false), or one lane per flagged 4-cell chunk running its four sums side by
// side (true: a quarter of the LDS reads, but a tile with few flagged chunks
// waits four times as long for its longest sums).
constexpr bool kChunkRecompute = false;

// One tile: the window [T, T + span) of row r is in x, visible to the block
// (la: a second buffer of the same size, stage: kPowStage entries per wave),
// err_e the filter's error terms (BoxErr::e, in LDS).  SEG: an interior
// window (span = 2048 + W - 1) goes through the segment filter, which takes
// the window's max |x| itself, with the limit coefficients lc (seg_coef, in
// LDS); every other window uses m, this thread's max |x| over its share of
// the window (+inf after a NaN).  Finds the tile's candidates and appends
// them to out.  Every thread of the block calls it; it ends with no barrier
// after its last LDS access, so a caller reusing x, la or stage synchronises
// first.
template <int QMAX, bool SEG>
__device__ __forceinline__ void pow2_tile(float* x, float* la, uint2* stage, float m, const float* row, int64_t T,
                                          int span, bool vec, int q, float thr, const double* err_e,
                                          const float2* lc, const float2* dvt, int64_t r, CandD* __restrict__ out,
                                          int64_t cap, unsigned long long* __restrict__ count) {
    const unsigned tid = opaque_tid();
    __shared__ uint32_t wbits[kThreads / 64];
    __shared__ float lim[kMaxPowQ + 1];
    const int W = 1 << q;
    const bool edge = span < kPowCells + W - 1;
    float M = 0.0f;
    uint32_t cf[2] = {0u, 0u};
    bool seg = false;
    if constexpr (SEG) {
        if (!edge) seg = pow2_flag_segments_at<0, QMAX>(q, x, la, lc, wbits, cf, M);  // block-uniform
    }
    if (!seg) {
        if (!SEG || edge) {
            m = __ockl_wfred_max_f32(m);  // wave max (m is never NaN: a NaN sample made it +inf)
            if ((tid & 63) == 0) wbits[tid >> 6] = __float_as_uint(m);
            __syncthreads();
            M = __uint_as_float(wbits[0]);
            for (int i = 1; i < kThreads / 64; ++i) M = fmaxf(M, __uint_as_float(wbits[i]));
        }
        // the per-cell paths' limits are block-uniform: thread k computes
        // level k's once (double-precision work that every wave would
        // otherwise repeat)
        if ((int)tid <= q) lim[tid] = level_limit(thr, (int)tid, err_e[tid], M);
        __syncthreads();
    }

    // Flag phase.  Flags are shifted in one bit per level, so level k ends at
    // bit q - k; reversing the word puts it at bit k.  The flagged cells are
    // then compacted into an LDS queue (cell << 14 | width flags; the level
    // buffers are free by then) in cell order, ranked by (slot, wave, lane)
    // with one ballot per slot; the exact recompute runs one chain per lane
    // over the queue (four, one per cell, for the segment filter's chunk
    // entries).
    uint32_t* queue = reinterpret_cast<uint32_t*>(la);
    __shared__ uint32_t slot_cnt[kPowPer][kThreads / 64];
    const int lane = tid & 63;
    const int wid = tid >> 6;
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    uint32_t total = 0;
    // Per-cell flags where exactness needs them: the last tile of a row (its
    // window is cut by the row end, so level spans are checked) and windows
    // with NaN/inf samples or magnitudes where a level sum could overflow.
    const bool careful = edge || !(M * (float)W < 0x1p125f);
    if (seg) {
        // segment layout: thread t holds chunks 2t and 2t + 1 (cells 8t ..
        // 8t + 7); a flagged chunk queues its four cells (or, with
        // kChunkRecompute, one entry for its first cell), ranked by (wave,
        // lane, chunk)
        cf[0] = __brev(cf[0]) >> (31 - q);
        cf[1] = __brev(cf[1]) >> (31 - q);
        const uint64_t b0 = __ballot(cf[0] != 0), b1 = __ballot(cf[1] != 0);
        if (lane == 0) slot_cnt[0][wid] = (uint32_t)(__popcll(b0) + __popcll(b1));
        __syncthreads();
        uint32_t pos = (uint32_t)(__popcll(b0 & below) + __popcll(b1 & below));
        for (int k = 0; k < kThreads / 64; ++k) {
            if (k < wid) pos += slot_cnt[0][k];
            total += slot_cnt[0][k];
        }
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            if (cf[s]) {
                const uint32_t c0 = 8u * tid + 4u * (uint32_t)s;
                if constexpr (kChunkRecompute) {
                    queue[pos] = (c0 << 14) | cf[s];
                } else {
                    reinterpret_cast<uint4*>(queue)[pos] =
                        make_uint4((c0 << 14) | cf[s], ((c0 + 1) << 14) | cf[s], ((c0 + 2) << 14) | cf[s],
                                   ((c0 + 3) << 14) | cf[s]);
                }
                ++pos;
            }
        }
        if constexpr (!kChunkRecompute) total *= 4;
    } else if (!SEG && !careful) {
        // chunk layout: slot s of thread t is chunk s*256 + t (4 cells); a
        // flagged chunk queues its four cells
        if constexpr (!SEG) pow2_flag_chunks_at<0, QMAX>(q, x, la, lim, cf);
        uint64_t mk[2];
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            cf[s] = __brev(cf[s]) >> (31 - q);
            mk[s] = __ballot(cf[s] != 0);
            if (lane == 0) slot_cnt[s][wid] = (uint32_t)__popcll(mk[s]);
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            uint32_t pos = total;
            for (int k = 0; k < kThreads / 64; ++k) {
                if (k < wid) pos += slot_cnt[s][k];
                total += slot_cnt[s][k];
            }
            if (cf[s]) {
                const uint32_t p = pos + (uint32_t)__popcll(mk[s] & below);
                const uint32_t c0 = 4u * (uint32_t)(s * kThreads + tid);
                reinterpret_cast<uint4*>(queue)[p] =
                    make_uint4((c0 << 14) | cf[s], ((c0 + 1) << 14) | cf[s], ((c0 + 2) << 14) | cf[s],
                               ((c0 + 3) << 14) | cf[s]);
            }
        }
        total *= 4;
    } else {
        // strided layout: slot j of thread t is cell j*256 + t
        uint32_t flags[kPowPer];
        if (span < kPowCells + W - 1)
            pow2_flag_rounds<true>(x, la, q, span, lim, flags);
        else
            pow2_flag_rounds<false>(x, la, q, span, lim, flags);
        uint64_t mk[kPowPer];
#pragma unroll
        for (int j = 0; j < kPowPer; ++j) {
            flags[j] = __brev(flags[j]) >> (31 - q);
            mk[j] = __ballot(flags[j] != 0);
            if (lane == 0) slot_cnt[j][wid] = (uint32_t)__popcll(mk[j]);
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kPowPer; ++j) {
            uint32_t pos = total;
            for (int k = 0; k < kThreads / 64; ++k) {
                if (k < wid) pos += slot_cnt[j][k];
                total += slot_cnt[j][k];
            }
            if (flags[j])
                queue[pos + (uint32_t)__popcll(mk[j] & below)] =
                    ((uint32_t)(j * kThreads + tid) << 14) | flags[j];
        }
    }
    if (total == 0) return;  // block-uniform
    // from q = 4 on the level rounds' second round stored its level over the
    // samples: reload them (the block's window was just read, so this mostly
    // hits L2); the segment filter leaves them in place
    if (q >= 4 && !seg) pow2_load_window(x, row + T, span, vec);
    __syncthreads();

    // Exact recompute: lane i of a round takes queue entry i and runs the
    // left-to-right running sum of its cell (of each of its chunk's four
    // cells after the segment filter) up to its largest flagged width.  The
    // queue is in cell order, so the lanes of a wave read neighbouring
    // samples (no LDS bank conflicts).  The loop bound is wave-uniform.
    // Candidates go to a per-wave LDS stage (cell << 4 | width index, S/N
    // bits).  A wave whose stage would overflow flushes it with one global
    // atomic; at the end the block adds its four stage fills and takes ONE
    // atomic for all of them (device-scope atomics on the shared counter
    // serialise, so their number, not the candidate count, sets the cost).
    const uint64_t lt_mask = below;
    const uint32_t wcap = kPowStage;
    uint2* wst = stage + (size_t)wid * wcap;
    uint32_t wn = 0;  // wave-uniform stage fill
    auto flush = [&]() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(count, (unsigned long long)wn);
        base = __shfl(base, 0);
        for (uint32_t i = (uint32_t)lane; i < wn; i += 64) {
            const uint2 v = wst[i];
            const unsigned long long g = base + i;
            if ((int64_t)g < cap)
                out[g] = CandD{(int32_t)(T + (v.x >> 4)), (int32_t)r, (int32_t)(v.x & 15u), __uint_as_float(v.y)};
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
        wn = 0;
    };
    // width 2^p of the cell at lt whose running sum covers 2^p samples
    // (wave-uniform call: the stage append is a wave operation)
    auto decide = [&](float sum, int lt, uint32_t fl, int p) {
        float snr = 0.0f;
        bool pred = false;
        if ((fl >> p) & 1u) {
            // dvt[p] = {sqrtf(2^p), 1 / it}: for even p the divisor is
            // 2^(p/2), so the IEEE quotient is the rounding of the exact
            // scaling sum * 2^(-p/2); for odd p the shared-reciprocal quotient
            const float2 dv = dvt[p];
            snr = (p & 1) ? div_rn(sum, dv.x, dv.y) : sum * dv.y;
            pred = snr > thr;
        }
        const uint64_t m = __ballot(pred);
        if (m) {
            const uint32_t n = (uint32_t)__popcll(m);
            if (wn + n > wcap) flush();
            if (pred)
                wst[wn + (uint32_t)__popcll(m & lt_mask)] =
                    make_uint2(((uint32_t)lt << 4) | (uint32_t)p, __float_as_uint(snr));
            wn += n;
        }
    };
    if (kChunkRecompute && seg) {
        // chunk entries (first cell c0, a multiple of 4): the four running
        // sums advance together, four samples per step from two aligned
        // float4s A = x[c0+4m .. c0+4m+3] and B = the next four; sum j adds
        // x[c0+j+4m .. c0+j+4m+3] in order.  Reads stay below 2304 floats:
        // c0 <= 2044 and the last step's B ends at c0 + 2^top + 3.
        for (uint32_t q0 = 0; q0 < total; q0 += kThreads) {  // block-uniform
            const uint32_t qi = q0 + tid;
            const uint32_t e = qi < total ? queue[qi] : 0u;
            const int c0 = (int)(e >> 14);
            const uint32_t fl = e & 0x3fffu;
            int top = fl ? 31 - __clz(fl) : -1;
            top = __builtin_amdgcn_readfirstlane(__ockl_wfred_max_i32(top));  // wave max: a scalar loop bound
            if (top < 0) continue;  // wave-uniform; no barrier in this loop
            const float4* x4 = reinterpret_cast<const float4*>(x) + (c0 >> 2);
            float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, s3 = 0.0f;
            auto decide4 = [&](int p) {
                decide(s0, c0, fl, p);
                decide(s1, c0 + 1, fl, p);
                decide(s2, c0 + 2, fl, p);
                decide(s3, c0 + 3, fl, p);
            };
            float4 A = x4[0], B = x4[1];
            s0 += A.x;
            s1 += A.y;
            s2 += A.z;
            s3 += A.w;
            decide4(0);
            if (top >= 1) {
                s0 += A.y;
                s1 += A.z;
                s2 += A.w;
                s3 += B.x;
                decide4(1);
            }
            if (top >= 2) {
                s0 += A.z;
                s1 += A.w;
                s2 += B.x;
                s3 += B.y;
                s0 += A.w;
                s1 += B.x;
                s2 += B.y;
                s3 += B.z;
                decide4(2);
            }
            if (top >= 3) {
                const int steps = 1 << (top - 2);
                float4 Bn = x4[2];
                for (int m = 1; m < steps; ++m) {  // wave-uniform
                    A = B;
                    B = Bn;
                    if (m + 2 <= steps) Bn = x4[m + 2];
                    s0 += A.x;
                    s1 += A.y;
                    s2 += A.z;
                    s3 += A.w;
                    s0 += A.y;
                    s1 += A.z;
                    s2 += A.w;
                    s3 += B.x;
                    s0 += A.z;
                    s1 += A.w;
                    s2 += B.x;
                    s3 += B.y;
                    s0 += A.w;
                    s1 += B.x;
                    s2 += B.y;
                    s3 += B.z;
                    const int w = 4 * (m + 1);
                    if ((w & (w - 1)) == 0) decide4(31 - __clz(w));
                }
            }
        }
    } else {
        for (uint32_t q0 = 0; q0 < total; q0 += kThreads) {  // block-uniform
            const uint32_t qi = q0 + tid;
            const uint32_t e = qi < total ? queue[qi] : 0u;
            const int lt = (int)(e >> 14);
            const uint32_t fl = e & 0x3fffu;
            int top = fl ? 31 - __clz(fl) : -1;
            top = __builtin_amdgcn_readfirstlane(__ockl_wfred_max_i32(top));  // wave max: a scalar loop bound
            if (top < 0) continue;  // wave-uniform; no barrier in this loop
            // The running sum s over samples lt, lt+1, ... is taken to the
            // wave's largest flagged width 2^top by every lane (uniform
            // control flow), and width 2^p is decided when s covers 2^p
            // samples.  A lane reads at most to lt + 2^top - 1 <= 2047 + W -
            // 1 < span_max; samples past its own flagged windows only enter
            // widths it does not decide.
            float s = 0.0f;
            const float* xs = x + lt;
            // widths 1 .. 16 (p <= 4): samples 0 .. 15
            {
                float v[16];
#pragma unroll
                for (int u = 0; u < 16; ++u) v[u] = (u < (1 << top)) ? xs[u] : 0.0f;
                s += v[0];
                decide(s, lt, fl, 0);
                if (top >= 1) {
                    s += v[1];
                    decide(s, lt, fl, 1);
                }
                if (top >= 2) {
                    s += v[2];
                    s += v[3];
                    decide(s, lt, fl, 2);
                }
                if (top >= 3) {
#pragma unroll
                    for (int u = 4; u < 8; ++u) s += v[u];
                    decide(s, lt, fl, 3);
                }
                if (top >= 4) {
#pragma unroll
                    for (int u = 8; u < 16; ++u) s += v[u];
                    decide(s, lt, fl, 4);
                }
            }
            // widths 32 .. 2^top: blocks of 16 samples, the next block's
            // reads issued before the current block is added
            if (top >= 5) {
                // two register blocks in turn (no copies): the block at b is
                // added while the one at b + 16 loads
                const int end = 1 << top;
                float A[16], B[16];
#pragma unroll
                for (int u = 0; u < 16; ++u) A[u] = xs[16 + u];
                for (int b = 16; b < end; b += 32) {  // wave-uniform
                    if (b + 16 < end) {
#pragma unroll
                        for (int u = 0; u < 16; ++u) B[u] = xs[b + 16 + u];
                    }
#pragma unroll
                    for (int u = 0; u < 16; ++u) s += A[u];
                    if (((b + 16) & (b + 15)) == 0) decide(s, lt, fl, 31 - __clz(b + 16));
                    if (b + 16 >= end) break;
                    if (b + 32 < end) {
#pragma unroll
                        for (int u = 0; u < 16; ++u) A[u] = xs[b + 32 + u];
                    }
#pragma unroll
                    for (int u = 0; u < 16; ++u) s += B[u];
                    if (((b + 32) & (b + 31)) == 0) decide(s, lt, fl, 31 - __clz(b + 32));
                }
            }
        }
    }
    // block-level final flush (the recompute loop is block-uniform, so every
    // thread reaches the barriers)
    __shared__ uint32_t wfill[kThreads / 64];
    __shared__ unsigned long long bbase;
    if (lane == 0) wfill[wid] = wn;
    __syncthreads();
    if (tid == 0) {
        uint32_t tot = 0;
        for (int k = 0; k < kThreads / 64; ++k) tot += wfill[k];
        bbase = tot ? atomicAdd(count, (unsigned long long)tot) : 0ull;
    }
    __syncthreads();
    unsigned long long base = bbase;
    for (int k = 0; k < wid; ++k) base += wfill[k];
    for (uint32_t i = (uint32_t)lane; i < wn; i += 64) {
        const uint2 v = wst[i];
        const unsigned long long g = base + i;
        if ((int64_t)g < cap)
            out[g] = CandD{(int32_t)(T + (v.x >> 4)), (int32_t)r, (int32_t)(v.x & 15u), __uint_as_float(v.y)};
    }
}

// One block = one 2048-cell tile of one row, any q <= kMaxPowQ (window in
// dynamic LDS: samples, level buffer, candidate stage).
__global__ void __launch_bounds__(kThreads) k_boxcar_pow2(const float* __restrict__ img, int64_t cols, int64_t pitch,
                                                          const float* __restrict__ d_thr, float h_thr, int q,
                                                          BoxErr err, CandD* __restrict__ out, int64_t cap,
                                                          unsigned long long* __restrict__ count) {
    extern __shared__ __attribute__((aligned(16))) float lds[];
    const int W = 1 << q;
    const int span_max = round_up4(kPowCells + W - 1);
    float* x = lds;                  // the samples, then every odd round's level
    float* la = lds + span_max;      // every even round's level, then the queue
    uint2* stage = reinterpret_cast<uint2*>(la + span_max);  // kPowStage per wave
    const int64_t r = blockIdx.y;
    const int64_t T = rotated_tile() * kPowCells;
    const int span = (int)min((int64_t)(kPowCells + W - 1), cols - T);
    const float* row = img + r * pitch;
    const bool vec = (pitch & 3) == 0 && (reinterpret_cast<uintptr_t>(img) & 15) == 0;
    __shared__ double err_e[kMaxPowQ + 1];
    __shared__ float2 dvt[kMaxPowQ + 1];
    if ((int)threadIdx.x <= q) {  // read after pow2_tile's first barrier
        err_e[threadIdx.x] = err.e[threadIdx.x];
        dvt[threadIdx.x] = box_divisor((int)threadIdx.x);
    }
    const float m = pow2_load_window(x, row + T, span, vec);
    pow2_tile<kMaxPowQ, false>(x, la, stage, m, row, T, span, vec, q, d_thr ? *d_thr : h_thr, err_e, nullptr, dvt, r,
                               out, cap, count);
}

// Windows up to q = QM in static LDS (QM = 8, or 9 for q = 9): two sample
// buffers, so that the next tile's window is fetched while the current one
// is searched.
constexpr int kPowPfTiles = 4;  // consecutive tiles of a row per block

// Starts copying floats [0, span4) of a window (span4 % 4 == 0, src 16-byte
// aligned, all of it inside the row) to LDS with global_load_lds_dwordx4:
// the memory unit writes lane l's 16 bytes at M0 + 16 l (M0: the wave's LDS
// byte address), so no registers are held while the loads are in flight.
// Issued as inline asm, so the compiler does not count these LDS writes:
// its barriers and LDS accesses then wait for nothing of them (the builtin
// would make every later barrier wait for the whole window).  The caller
// waits itself: complete once this wave's vmcnt is zero, visible to the
// block after a barrier that follows, and nothing reads or writes dst
// before then.
__device__ __forceinline__ void pow2_issue_window(float* dst, const float* src, int span4) {
    const unsigned tid = opaque_tid();
    const int n4 = span4 >> 2;
    // the wave's first float4 and its LDS address are wave-uniform (SGPRs);
    // a lane adds only its 32-bit byte offset (the saddr form of the load)
    const int wbase = (int)__builtin_amdgcn_readfirstlane(tid & ~63u);
    const uint32_t voff = 16u * (tid & 63u);
    for (int j0 = 0; j0 < n4; j0 += kThreads) {  // block-uniform
        const int i = j0 + (int)tid;
        const uint32_t m0 = __builtin_amdgcn_readfirstlane(
            (uint32_t)(uintptr_t)(__attribute__((address_space(3))) float*)(dst + 4 * (j0 + wbase)));
        const float* g = src + 4 * (j0 + wbase);
        // M0 is saved and restored around the load; one wait state between
        // the SALU write of M0 and the load that reads it
        uint32_t saved;
        if (i < n4)
            asm volatile("s_mov_b32 %0, m0\n\ts_mov_b32 m0, %1\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %2, %3\n\t"
                         "s_mov_b32 m0, %0"
                         : "=&s"(saved)
                         : "s"(m0), "v"(voff), "s"(g));
    }
}

// One block = kPowPfTiles consecutive tiles of one row (q <= QM).  A
// tile whose whole rounded window lies inside a 16-byte aligned row is
// fetched by global_load_lds, issued when the previous tile starts; the row's
// last tiles are loaded directly.
template <int QM>
__global__ void __launch_bounds__(kThreads) k_boxcar_pow2_pf(const float* __restrict__ img, int64_t cols,
                                                             int64_t pitch, const float* __restrict__ d_thr,
                                                             float h_thr, int q, BoxErr err,
                                                             CandD* __restrict__ out, int64_t cap,
                                                             unsigned long long* __restrict__ count) {
    static_assert(QM <= kPowPfMaxQ, "segment filter window");
    constexpr int kSpan = round_up4(kPowCells + (1 << QM) - 1);
    __shared__ __attribute__((aligned(16))) float xa[kSpan];
    __shared__ __attribute__((aligned(16))) float xb[kSpan];
    __shared__ __attribute__((aligned(16))) float la[kSpan];
    __shared__ __attribute__((aligned(16))) uint2 stage[kPowStage * (kThreads / 64)];
    __shared__ double err_e[QM + 1];
    __shared__ float2 lc[QM + 1];
    __shared__ float2 dvt[QM + 1];
    const float thr = d_thr ? *d_thr : h_thr;
    if ((int)threadIdx.x <= q) {  // read after the first tile's barriers
        err_e[threadIdx.x] = err.e[threadIdx.x];
        lc[threadIdx.x] = seg_coef(thr, (int)threadIdx.x, err.e[threadIdx.x]);
        dvt[threadIdx.x] = box_divisor((int)threadIdx.x);
    }
    const int full = kPowCells + (1 << q) - 1;
    const int full4 = round_up4(full);
    const int64_t r = blockIdx.y;
    const int64_t ntiles = ceil_div(cols, (int64_t)kPowCells);
    const int64_t k0 = rotated_tile() * kPowPfTiles;
    const int64_t k1 = min(ntiles, k0 + (int64_t)kPowPfTiles);
    const float* row = img + r * pitch;
    const bool vec = (pitch & 3) == 0 && (reinterpret_cast<uintptr_t>(img) & 15) == 0;
    auto dma = [&](int64_t k) { return vec && k * kPowCells + full4 <= cols; };  // monotone in k
    auto step = [&](float* cur, float* nxt, int64_t k) {
        const int64_t T = k * kPowCells;
        const int span = (int)min((int64_t)full, cols - T);
        float m = 0.0f;  // a DMA'd window is interior: the segment filter takes its max itself
        if (dma(k)) {
            asm volatile("s_waitcnt vmcnt(0)" : : : "memory");  // this wave's part of the window has landed
            __syncthreads();
            if (k + 1 < k1 && dma(k + 1)) pow2_issue_window(nxt, row + T + kPowCells, full4);
        } else {
            __syncthreads();
            m = pow2_load_window(cur, row + T, span, vec);
            __syncthreads();
        }
        pow2_tile<QM, true>(cur, la, stage, m, row, T, span, vec, q, thr, err_e, lc, dvt, r, out, cap, count);
    };
    if (k0 < k1 && dma(k0)) pow2_issue_window(xa, row + k0 * kPowCells, full4);
    for (int64_t k = k0; k < k1; ++k) {  // block-uniform
        const bool odd = ((k - k0) & 1) != 0;
        step(odd ? xb : xa, odd ? xa : xb, k);
    }
}

// digitize_kernel (Clusterization.cu:206-214): x = floorf(x) in place
__global__ void __launch_bounds__(kThreads) k_digitize(float* __restrict__ x, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
        x[i] = floorf(x[i]);
}

__global__ void __launch_bounds__(kThreads) k_boxcar_snr(const float* __restrict__ img, int64_t cols, int w,
                                                         float* __restrict__ snr) {
    const int64_t r = blockIdx.y;
    const int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (t >= cols) return;
    float v = 0.0f;
    if (t + w <= cols) {
        float s = 0.0f;
        for (int j = 0; j < w; ++j) s += img[r * cols + t + j];
        v = s / sqrtf((float)w);
    }
    snr[r * cols + t] = v;
}

__global__ void __launch_bounds__(kThreads) k_boxcar_snr_adjoint(const float* __restrict__ g, int64_t cols, int w,
                                                                 float* __restrict__ gi) {
    const int64_t r = blockIdx.y;
    const int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (t >= cols) return;
    const int64_t lo = t - w + 1 > 0 ? t - w + 1 : 0;
    const int64_t hi = t < cols - w ? t : cols - w;
    float s = 0.0f;
    for (int64_t k = lo; k <= hi; ++k) s += g[r * cols + k];
    gi[r * cols + t] = s / sqrtf((float)w);
}

// ---------------------------------------------------------------------------
// single-maximum detector (Best, better, detect_windows, k_detect_final in
// detect.hpp)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kThreads) k_detect_blocks(const float* __restrict__ a, const float* __restrict__ n,
                                                            int64_t cols, int wnd, Best* __restrict__ blocks) {
    __shared__ Best sh[kThreads];
    const int64_t r = blockIdx.y;
    const int64_t t = (int64_t)blockIdx.x * kDetCells + threadIdx.x;
    Best b{0.0f, -1, -1, -1};
    if (t < cols) detect_windows(a, n, r, t, cols, wnd, b);
    block_best(b, sh);
    if (threadIdx.x == 0) blocks[(int64_t)blockIdx.y * gridDim.x + blockIdx.x] = sh[0];
}

// ---------------------------------------------------------------------------
// tiled single-maximum detectors with the window arithmetic of the
// reference's per-block kernels (detection/kernel.cu:42-226): a row is cut
// into tiles of `tile` columns (the caller's block width); inside a tile the
// values are prefix sums accumulated sequentially in fp32 from the tile's
// first column, and a window's value is a difference of two prefix sums.
//   mode 1 (multi_windowing_kernel_v1, :42-133): y = x / sqrtf(1e-7 + n),
//     value = (Y[t+w] - Y[t]) / sqrtf(w); the window lies inside its tile.
//   mode 2 (multi_windowing_kernel_v2, :136-226): value =
//     sqrtf((X2[t+w] - X2[t]) / (N2[t+w] - N2[t])) with X2, N2 the prefix sums
//     of x^2 and n^2 over the tile and the W samples after it, and N2[0] =
//     1e-6 as the reference initialises it.
// Windows must end inside the row.  Per tile the best window (largest value;
// ties to the smaller width, then column) goes to the caller's per-tile
// arrays {value, row*cols + column, width}; a second kernel reduces them
// (ties to the smaller width, then row, then column).
// ---------------------------------------------------------------------------
constexpr int kDetWave = 64;

__device__ __forceinline__ bool det_better(float v, int w, int64_t pos, float bv, int bw, int64_t bpos) {
    if (!(v == v)) return false;  // NaN never wins
    if (bw < 0 || !(bv == bv)) return true;
    if (v != bv) return v > bv;
    if (w != bw) return w < bw;
    return pos < bpos;
}

template <int MODE>
__global__ void __launch_bounds__(kDetWave) k_detect_tiles(const float* __restrict__ x, const float* __restrict__ n,
                                                           int64_t cols, int tile, int wnd,
                                                           float* __restrict__ tile_val,
                                                           int32_t* __restrict__ tile_pos,
                                                           int32_t* __restrict__ tile_w) {
    extern __shared__ float det_lds[];
    const int64_t r = blockIdx.y;
    const int64_t c0 = (int64_t)blockIdx.x * tile;
    // samples this tile's prefix sums cover
    const int64_t span = MODE == 1 ? tile : (int64_t)tile + wnd;
    const int64_t navail = cols - c0 < span ? cols - c0 : span;
    float* A = det_lds;             // per-sample y (mode 1) or x^2 (mode 2), then prefix sums
    float* B = det_lds + span + 1;  // n^2 then its prefix sums (mode 2)
    const float* xr = x + r * cols + c0;
    const float* nr = n ? n + r * cols + c0 : nullptr;
    for (int64_t i = threadIdx.x; i < navail; i += kDetWave) {
        const float xv = xr[i];
        if (MODE == 1) {
            A[i + 1] = nr ? xv / sqrtf((float)(1.0E-7 + (double)nr[i])) : xv;
        } else {
            const float nv = nr ? nr[i] : 1.0f;
            A[i + 1] = xv * xv;
            B[i + 1] = nv * nv;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float s = 0.0f, q = 0.0f;
        A[0] = 0.0f;
        if (MODE == 2) B[0] = 0.000001f;
        for (int64_t i = 0; i < navail; ++i) {
            s = s + A[i + 1];
            A[i + 1] = s;
            if (MODE == 2) {
                q = q + B[i + 1];
                B[i + 1] = q;
            }
        }
    }
    __syncthreads();
    float bv = 0.0f;
    int bw = -1;
    int64_t bpos = 0;
    const int64_t tmax = cols - c0 < tile ? cols - c0 : tile;
    for (int64_t t = threadIdx.x; t < tmax; t += kDetWave) {
        for (int w = 1; w <= wnd; ++w) {
            if (t + w > navail) break;                  // past the row's end (or, mode 1, the tile's)
            float v;
            if (MODE == 1)
                v = (A[t + w] - A[t]) / sqrtf((float)w);
            else
                v = sqrtf((A[t + w] - A[t]) / (B[t + w] - B[t]));
            if (det_better(v, w, t, bv, bw, bpos)) {
                bv = v;
                bw = w;
                bpos = t;
            }
        }
    }
    // wave reduction: (value desc, width asc, column asc)
    for (int off = kDetWave / 2; off > 0; off >>= 1) {
        const float ov = __shfl_down(bv, off);
        const int ow = __shfl_down(bw, off);
        const int64_t op = __shfl_down(bpos, off);
        if (ow >= 0 && det_better(ov, ow, op, bv, bw, bpos)) {
            bv = ov;
            bw = ow;
            bpos = op;
        }
    }
    if (threadIdx.x == 0) {
        const int64_t k = (int64_t)blockIdx.y * gridDim.x + blockIdx.x;
        tile_val[k] = bv;
        tile_pos[k] = bw >= 0 ? (int32_t)(r * cols + c0 + bpos) : -1;
        tile_w[k] = bw;
    }
}

__global__ void __launch_bounds__(kThreads) k_detect_tiles_final(const float* __restrict__ tile_val,
                                                                 const int32_t* __restrict__ tile_pos,
                                                                 const int32_t* __restrict__ tile_w, int64_t ntiles,
                                                                 int64_t cols, sps_detection* __restrict__ out) {
    __shared__ Best sh[kThreads];
    Best b{0.0f, -1, -1, -1};
    for (int64_t i = threadIdx.x; i < ntiles; i += kThreads) {
        if (tile_w[i] < 0) continue;
        Best c{tile_val[i], tile_w[i], (int32_t)(tile_pos[i] / cols), (int32_t)(tile_pos[i] % cols)};
        if (c.snr == c.snr && better(c, b)) b = c;
    }
    block_best(b, sh);
    if (threadIdx.x == 0) {
        out->iwidth = sh[0].w;
        out->irow = sh[0].r;
        out->icol = sh[0].t;
        out->snr = sh[0].snr;
    }
}

// ---------------------------------------------------------------------------
// row cleaning without normalisation (clean_fdmtInput_and_create_normInput,
// CDMT/Statistical_preprocessing.cu:102-120,142-268): rows whose variance is
// more than 4 sigma from the mean row variance are zeroed; the 0/1 mask is a
// byte per sample; buff = {row variances[R], mean variance, std of variances}.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kThreads) k_clean_rows(float* __restrict__ x, uint8_t* __restrict__ mask,
                                                         int64_t cols, const uint8_t* __restrict__ keep) {
    const int64_t r = blockIdx.y;
    const bool k = keep[r] != 0;
    for (int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x; t < cols; t += (int64_t)gridDim.x * kThreads) {
        const int64_t i = r * cols + t;
        if (!k) x[i] = 0.0f;
        if (mask) mask[i] = k ? 1 : 0;
    }
}

__global__ void __launch_bounds__(kThreads) k_clean_buff(const double* __restrict__ rv, const double* __restrict__ stats,
                                                         int64_t rows, float* __restrict__ buff) {
    for (int64_t r = threadIdx.x; r < rows; r += kThreads) buff[r] = (float)rv[r];
    if (threadIdx.x == 0) {
        buff[rows] = (float)stats[2];
        buff[rows + 1] = (float)stats[3];
    }
}

}  // namespace

// ---------------------------------------------------------------------------
void launch_row_stats(const void* in, int in_u8, int64_t rows, int64_t cols, double* d_row_mean, double* d_row_var,
                      hipStream_t st) {
    if (rows <= 0 || cols <= 0) return;
    if (rows > 0x7fffffffLL) throw Error(SPS_ERR_INVALID_ARGUMENT, "too many rows");
    if (in_u8)
        hipLaunchKernelGGL((k_row_stats<true>), dim3((uint32_t)rows), dim3(kLanes), 0, st, in, cols, d_row_mean,
                           d_row_var);
    else
        hipLaunchKernelGGL((k_row_stats<false>), dim3((uint32_t)rows), dim3(kLanes), 0, st, in, cols, d_row_mean,
                           d_row_var);
    SPS_HIP(hipGetLastError());
}

void launch_stats_finalize(const double* d_row_mean, const double* d_row_var, int64_t rows, int mode, double* d_stats,
                           float* d_mu, float* d_sigma, uint8_t* d_keep, float* d_buff, hipStream_t st) {
    hipLaunchKernelGGL(k_stats_finalize, dim3(1), dim3(kLanes), 0, st, d_row_mean, d_row_var, rows, mode, d_stats,
                       d_mu, d_sigma, d_keep, d_buff);
    SPS_HIP(hipGetLastError());
}

static dim3 row_grid(int64_t rows, int64_t cols) {
    int64_t bx = ceil_div(cols, kThreads);
    if (bx > 1024) bx = 1024;
    return dim3((uint32_t)bx, (uint32_t)rows);
}

void launch_apply_norm(float* x, float* mask, int64_t rows, int64_t cols, const float* d_mu, const float* d_sigma,
                       const uint8_t* d_keep, hipStream_t st) {
    if (rows <= 0 || cols <= 0) return;
    if (rows > 65535) throw Error(SPS_ERR_INVALID_ARGUMENT, "more than 65535 rows");
    hipLaunchKernelGGL(k_apply_norm, row_grid(rows, cols), dim3(kThreads), 0, st, x, mask, cols, d_mu, d_sigma,
                       d_keep);
    SPS_HIP(hipGetLastError());
}

void launch_fdmt_normalization_rows(const float* a, int64_t rows, int64_t cols, const float* n, int64_t norm_cols,
                                    float* out, hipStream_t st) {
    if (rows <= 0 || cols <= 0) return;
    if (rows > 65535) throw Error(SPS_ERR_INVALID_ARGUMENT, "more than 65535 rows");
    hipLaunchKernelGGL(k_fdmt_normalization_rows, row_grid(rows, cols), dim3(kThreads), 0, st, a, cols, n, norm_cols,
                       out);
    SPS_HIP(hipGetLastError());
}

void launch_fdmt_normalization(const float* a, const float* n, int64_t len, float* out, hipStream_t st) {
    if (len <= 0) return;
    int64_t nb = ceil_div(len, kThreads);
    if (nb > 65536) nb = 65536;
    hipLaunchKernelGGL(k_fdmt_normalization, dim3((uint32_t)nb), dim3(kThreads), 0, st, a, n, len, out);
    SPS_HIP(hipGetLastError());
}

// One block per output row: the histogram of the kept channels' shifts in
// LDS, its inclusive prefix sum (per-thread contiguous chunks, then a scan of
// the chunk totals), written as floats.  Counts are integers below 2^24: exact.
__global__ void __launch_bounds__(kThreads) k_mask_counts(const int16_t* __restrict__ shift,
                                                          const uint8_t* __restrict__ keep, int64_t nchans, int L,
                                                          float* __restrict__ out, int64_t pitch) {
    extern __shared__ int hist[];  // L counters, then kThreads chunk totals
    int* tot = hist + L;
    const int64_t r = blockIdx.x;
    for (int t = threadIdx.x; t < L; t += kThreads) hist[t] = 0;
    __syncthreads();
    const int16_t* sr = shift + r * nchans;
    for (int64_t c = threadIdx.x; c < nchans; c += kThreads) {
        const int s = sr[c];
        if (keep[c] && s >= 0 && s < L) atomicAdd(&hist[s], 1);
    }
    __syncthreads();
    const int chunk = (L + kThreads - 1) / kThreads;
    const int t0 = threadIdx.x * chunk;
    const int t1 = min(L, t0 + chunk);
    int acc = 0;
    for (int t = t0; t < t1; ++t) {
        acc += hist[t];
        hist[t] = acc;
    }
    // inclusive scan of the chunk totals (Hillis-Steele), then exclusive
    int v = acc;
    for (int off = 1; off < kThreads; off <<= 1) {
        __syncthreads();
        tot[threadIdx.x] = v;
        __syncthreads();
        if ((int)threadIdx.x >= off) v += tot[threadIdx.x - off];
    }
    const int base = v - acc;
    for (int t = t0; t < t1; ++t) hist[t] += base;
    __syncthreads();
    float* o = out + r * pitch;
    for (int t = threadIdx.x; t < L; t += kThreads) o[t] = (float)hist[t];
}

void launch_mask_counts(const int16_t* shift, const uint8_t* keep, int64_t rows, int64_t nchans, int L, float* out,
                        int64_t pitch, hipStream_t st) {
    if (rows <= 0 || L <= 0) return;
    const size_t lds = sizeof(int) * ((size_t)L + kThreads);
    hipLaunchKernelGGL(k_mask_counts, dim3((uint32_t)rows), dim3(kThreads), lds, st, shift, keep, nchans, L, out,
                       pitch);
    SPS_HIP(hipGetLastError());
}

void launch_snr_normalise(float* img, int64_t rows, int64_t cols, int64_t pitch, const float* d_norm,
                          int64_t norm_len, const double* d_stats, hipStream_t st) {
    if (rows <= 0 || cols <= 0) return;
    if (rows > 65535) throw Error(SPS_ERR_INVALID_ARGUMENT, "more than 65535 rows");
    hipLaunchKernelGGL(k_snr_normalise, row_grid(rows, cols), dim3(kThreads), 0, st, img, cols, pitch, d_norm,
                       norm_len, d_stats);
    SPS_HIP(hipGetLastError());
}

void launch_boxcar(const float* img, int64_t rows, int64_t cols, int64_t pitch, const float* d_thr, float h_thr,
                   int wmax, int pow2, CandD* out, int64_t cap, unsigned long long* d_count, hipStream_t st) {
    if (rows <= 0 || cols <= 0 || wmax <= 0) return;
    if (rows > 65535) throw Error(SPS_ERR_INVALID_ARGUMENT, "more than 65535 rows");
    if (pow2) {
        int q = 0;
        while ((1 << q) < wmax) ++q;
        if ((1 << q) != wmax || q > kMaxPowQ) throw Error(SPS_ERR_INVALID_ARGUMENT, "pow2 window must be 2^q, q <= 13");
        const dim3 grid((uint32_t)ceil_div(cols, kPowCells), (uint32_t)rows);
        const size_t lds = sizeof(float) * 2 * (size_t)round_up4(kPowCells + wmax - 1) +
                           sizeof(uint2) * (size_t)kPowStage * (kThreads / 64);
        if (lds > 150 * 1024) throw Error(SPS_ERR_INVALID_ARGUMENT, "boxcar window too wide for LDS");
        // error terms of the filter: (gamma_{w-1} + gamma_k) * w * (1 + 2^-20)
        BoxErr err{};
        const double u = 0x1p-24;
        auto gam = [&](double n) { return n * u / (1.0 - n * u); };
        for (int k = 0; k <= kMaxPowQ; ++k) {
            const double w = std::ldexp(1.0, k);
            // pairwise level sums: gamma_k w; the segment bound U of widths
            // >= 16 (segment sums of 8 under a depth k-3 tree, two prefix
            // extremes of <= 7 terms, two more roundings): gamma_{k+6} (w + 16)
            const double ours = k >= 4 ? gam((double)k + 6.0) * (w + 16.0) : gam((double)k) * w;
            err.e[k] = (gam(w - 1.0) * w + ours) * (1.0 + 0x1p-20);
        }
        if (q <= kPowPfMaxQ) {
            // static LDS sized for q <= 8 (5 blocks per CU), or for q = 9
            const dim3 gpf((uint32_t)ceil_div(ceil_div(cols, kPowCells), (int64_t)kPowPfTiles), (uint32_t)rows);
            if (q <= 8)
                hipLaunchKernelGGL(k_boxcar_pow2_pf<8>, gpf, dim3(kThreads), 0, st, img, cols, pitch, d_thr, h_thr, q,
                                   err, out, cap, d_count);
            else
                hipLaunchKernelGGL(k_boxcar_pow2_pf<kPowPfMaxQ>, gpf, dim3(kThreads), 0, st, img, cols, pitch, d_thr,
                                   h_thr, q, err, out, cap, d_count);
        } else {
            hipLaunchKernelGGL(k_boxcar_pow2, grid, dim3(kThreads), lds, st, img, cols, pitch, d_thr, h_thr, q, err,
                               out, cap, d_count);
        }
    } else {
        if (wmax > 8192) throw Error(SPS_ERR_INVALID_ARGUMENT, "window wider than 8192 samples");
        const dim3 grid((uint32_t)ceil_div(cols, kBoxCells), (uint32_t)rows);
        const size_t lds = sizeof(float) * (size_t)(kBoxCells + wmax);
        hipLaunchKernelGGL(k_boxcar_all, grid, dim3(kThreads), lds, st, img, cols, pitch, d_thr, h_thr, wmax, out, cap,
                           d_count);
    }
    SPS_HIP(hipGetLastError());
}

void launch_boxcar_snr(const float* img, int64_t rows, int64_t cols, int w, float* snr, hipStream_t st) {
    if (rows <= 0 || cols <= 0) return;
    if (rows > 65535) throw Error(SPS_ERR_INVALID_ARGUMENT, "more than 65535 rows");
    hipLaunchKernelGGL(k_boxcar_snr, dim3((uint32_t)ceil_div(cols, kThreads), (uint32_t)rows), dim3(kThreads), 0, st,
                       img, cols, w, snr);
    SPS_HIP(hipGetLastError());
}

void launch_digitize(float* x, int64_t n, hipStream_t st) {
    if (n <= 0) return;
    const int64_t nb = std::min<int64_t>(ceil_div(n, kThreads), 65536);
    hipLaunchKernelGGL(k_digitize, dim3((uint32_t)nb), dim3(kThreads), 0, st, x, n);
    SPS_HIP(hipGetLastError());
}

void launch_boxcar_snr_adjoint(const float* g, int64_t rows, int64_t cols, int w, float* gi, hipStream_t st) {
    if (rows <= 0 || cols <= 0) return;
    if (rows > 65535) throw Error(SPS_ERR_INVALID_ARGUMENT, "more than 65535 rows");
    hipLaunchKernelGGL(k_boxcar_snr_adjoint, dim3((uint32_t)ceil_div(cols, kThreads), (uint32_t)rows), dim3(kThreads),
                       0, st, g, cols, w, gi);
    SPS_HIP(hipGetLastError());
}

int64_t detect_max_blocks(int64_t rows, int64_t cols) { return ceil_div(cols, kDetCells) * rows; }

void launch_detect_tiles(int mode, const float* x, const float* n, int64_t rows, int64_t cols, int tile, int wnd,
                         float* tile_val, int32_t* tile_pos, int32_t* tile_w, void* d_out, hipStream_t st) {
    if (mode != 1 && mode != 2) throw Error(SPS_ERR_INVALID_ARGUMENT, "detector mode must be 1 or 2");
    if (tile < 1 || tile > 4096 || wnd < 1 || wnd > 4096) throw Error(SPS_ERR_INVALID_ARGUMENT, "tile and window must be 1..4096");
    if (rows < 1 || rows > 65535 || cols < 1) throw Error(SPS_ERR_INVALID_ARGUMENT, "rows must be 1..65535");
    if (cols > 0x7fffffffLL / rows)
        throw Error(SPS_ERR_INVALID_ARGUMENT, "the per-tile index row*cols+column is a 32-bit int");
    const int64_t nt = ceil_div(cols, tile);
    if (nt > 0x7fffffffLL) throw Error(SPS_ERR_INVALID_ARGUMENT, "too many tiles");
    const int64_t span = mode == 1 ? tile : (int64_t)tile + wnd;
    const size_t lds = sizeof(float) * (size_t)(2 * (span + 1));
    const dim3 grid((uint32_t)nt, (uint32_t)rows);
    if (mode == 1)
        hipLaunchKernelGGL(k_detect_tiles<1>, grid, dim3(kDetWave), lds, st, x, n, cols, tile, wnd, tile_val, tile_pos,
                           tile_w);
    else
        hipLaunchKernelGGL(k_detect_tiles<2>, grid, dim3(kDetWave), lds, st, x, n, cols, tile, wnd, tile_val, tile_pos,
                           tile_w);
    hipLaunchKernelGGL(k_detect_tiles_final, dim3(1), dim3(kThreads), 0, st, tile_val, tile_pos, tile_w, nt * rows,
                       cols, static_cast<sps_detection*>(d_out));
    SPS_HIP(hipGetLastError());
}

void launch_clean_rows(float* x, uint8_t* mask, int64_t rows, int64_t cols, const uint8_t* d_keep,
                       const double* d_row_var, const double* d_stats, float* d_buff, hipStream_t st) {
    if (rows <= 0 || cols <= 0) return;
    if (rows > 65535) throw Error(SPS_ERR_INVALID_ARGUMENT, "more than 65535 rows");
    hipLaunchKernelGGL(k_clean_rows, row_grid(rows, cols), dim3(kThreads), 0, st, x, mask, cols, d_keep);
    if (d_buff) hipLaunchKernelGGL(k_clean_buff, dim3(1), dim3(kThreads), 0, st, d_row_var, d_stats, rows, d_buff);
    SPS_HIP(hipGetLastError());
}

void launch_detect_max(const float* a, const float* n, int64_t rows, int64_t cols, int wnd, void* d_block_best,
                       int64_t nblocks_cap, void* d_out, hipStream_t st) {
    const int64_t nb = detect_max_blocks(rows, cols);
    if (nb > nblocks_cap) throw Error(SPS_ERR_INVALID_ARGUMENT, "detector scratch too small");
    if (rows > 65535) throw Error(SPS_ERR_INVALID_ARGUMENT, "more than 65535 rows");
    hipLaunchKernelGGL(k_detect_blocks, dim3((uint32_t)ceil_div(cols, kDetCells), (uint32_t)rows), dim3(kThreads), 0,
                       st, a, n, cols, wnd, static_cast<Best*>(d_block_best));
    hipLaunchKernelGGL(k_detect_final, dim3(1), dim3(kThreads), 0, st, static_cast<const Best*>(d_block_best), nb,
                       static_cast<sps_detection*>(d_out));
    SPS_HIP(hipGetLastError());
}

}  // namespace sps
