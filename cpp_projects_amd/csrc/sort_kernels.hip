// sort_kernels.hip -- stable LSD radix sort, scans and candidate clustering.
//
// The clustering of Clusterization.cu:92-166,793-998 (per-group segmented
// sorts by mt, mdt, mwidth with gap splits) is expressed as three global
// stable sorts on composite keys (group id << 32 | member), which visit the
// same groups in the same member order as a stable segmented sort.
#include "common.hpp"
#include "kernels.hpp"

namespace sps {

namespace {

constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;
constexpr int kItemsPerThread = 16;
constexpr int kSortBlock = kThreads * kItemsPerThread;  // 4096 items per block
constexpr int kWave = 64;
constexpr int kWaves = kThreads / kWave;

__global__ void __launch_bounds__(kThreads) k_radix_hist(const uint64_t* __restrict__ keys, int64_t n, int shift,
                                                         uint32_t* __restrict__ hist, int64_t nblocks) {
    __shared__ uint32_t h[kRadix];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t b0 = (int64_t)blockIdx.x * kSortBlock;
    for (int j = 0; j < kItemsPerThread; ++j) {
        const int64_t i = b0 + (int64_t)j * kThreads + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & (kRadix - 1)], 1u);
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Items of a block are ranked in index order: chunk by chunk (256 items, one
// per thread), wave peers with equal digits found with 8 ballots, waves
// combined through LDS counts.  Equal digits keep their input order.
__global__ void __launch_bounds__(kThreads) k_radix_scatter(const uint64_t* __restrict__ keys,
                                                            const uint32_t* __restrict__ vals, int64_t n, int shift,
                                                            const uint32_t* __restrict__ offs, int64_t nblocks,
                                                            uint64_t* __restrict__ keys_out,
                                                            uint32_t* __restrict__ vals_out) {
    __shared__ uint32_t running[kRadix];
    __shared__ uint32_t wcount[kWaves][kRadix];
    const int lane = threadIdx.x % kWave;
    const int wave = threadIdx.x / kWave;
    running[threadIdx.x] = offs[(int64_t)threadIdx.x * nblocks + blockIdx.x];
    for (int w = 0; w < kWaves; ++w) wcount[w][threadIdx.x] = 0;
    __syncthreads();
    const int64_t b0 = (int64_t)blockIdx.x * kSortBlock;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int j = 0; j < kItemsPerThread; ++j) {
        const int64_t i = b0 + (int64_t)j * kThreads + threadIdx.x;
        const bool valid = i < n;
        uint64_t k = 0;
        uint32_t v = 0;
        uint32_t dg = 0;
        if (valid) {
            k = keys[i];
            v = vals[i];
            dg = (uint32_t)((k >> shift) & (kRadix - 1));
        }
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < kRadixBits; ++bit) {
            const uint64_t bal = __ballot(valid && ((dg >> bit) & 1u));
            peers &= ((dg >> bit) & 1u) ? bal : ~bal;
        }
        uint32_t rank = 0;
        if (valid) {
            rank = (uint32_t)__popcll(peers & lt_mask);
            const uint32_t cnt = (uint32_t)__popcll(peers);
            if (rank == 0) wcount[wave][dg] = cnt;
        }
        __syncthreads();
        if (valid) {
            uint32_t pre = 0;
            for (int w = 0; w < wave; ++w) pre += wcount[w][dg];
            const uint32_t pos = running[dg] + pre + rank;
            keys_out[pos] = k;
            vals_out[pos] = v;
        }
        __syncthreads();
        uint32_t tot = 0;
        for (int w = 0; w < kWaves; ++w) {
            tot += wcount[w][threadIdx.x];
            wcount[w][threadIdx.x] = 0;
        }
        running[threadIdx.x] += tot;
        __syncthreads();
    }
}

// ---- scan: inclusive, uint32, blocks of 1024 items (4 per thread) ----
constexpr int kScanItems = 4;
constexpr int kScanBlock = kThreads * kScanItems;

// exclusive prefix of v over the block (Hillis-Steele over the thread
// values; the exclusive value is the previous thread's inclusive one, so no
// subtraction enters a float scan)
template <typename T>
__device__ T block_excl_scan(T v, T* sh, T* total) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < kThreads; off <<= 1) {
        const T add = (int)threadIdx.x >= off ? sh[threadIdx.x - off] : T(0);
        __syncthreads();
        sh[threadIdx.x] += add;
        __syncthreads();
    }
    const T excl = threadIdx.x > 0 ? sh[threadIdx.x - 1] : T(0);
    *total = sh[kThreads - 1];
    __syncthreads();
    return excl;
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_scan_local(T* __restrict__ data, int64_t n, T* __restrict__ block_sums) {
    __shared__ T sh[kThreads];
    const int64_t base = (int64_t)blockIdx.x * kScanBlock + (int64_t)threadIdx.x * kScanItems;
    T x[kScanItems];
    T s = T(0);
    for (int e = 0; e < kScanItems; ++e) {
        x[e] = (base + e < n) ? data[base + e] : T(0);
        s += x[e];
    }
    T total;
    T pre = block_excl_scan<T>(s, sh, &total);
    for (int e = 0; e < kScanItems; ++e) {
        pre += x[e];
        if (base + e < n) data[base + e] = pre;
    }
    if (threadIdx.x == 0 && block_sums) block_sums[blockIdx.x] = total;
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_scan_add(T* __restrict__ data, int64_t n,
                                                       const T* __restrict__ block_prefix) {
    if (blockIdx.x == 0) return;
    const T add = block_prefix[blockIdx.x - 1];
    const int64_t base = (int64_t)blockIdx.x * kScanBlock;
    for (int e = threadIdx.x; e < kScanBlock; e += kThreads)
        if (base + e < n) data[base + e] += add;
}

__global__ void k_excl_from_incl(const uint32_t* __restrict__ incl, uint32_t* __restrict__ excl, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i < n) excl[i] = (i == 0) ? 0u : incl[i - 1];
}

// ---- candidates ----
__device__ __forceinline__ uint32_t ord_i32(int32_t v) { return (uint32_t)v ^ 0x80000000u; }
__device__ __forceinline__ int32_t from_ord(uint32_t u) { return (int32_t)(u ^ 0x80000000u); }
__device__ __forceinline__ uint32_t ord_f32(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__global__ void k_cand_keys(const CandD* __restrict__ c, int64_t n, int bits_t, int bits_w,
                            uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const CandD x = c[i];
    keys[i] = ((uint64_t)(uint32_t)x.mdt << (bits_t + bits_w)) | ((uint64_t)(uint32_t)x.mt << bits_w) |
              (uint64_t)(uint32_t)x.mwidth;
    vals[i] = __float_as_uint(x.msnr);
}

__global__ void k_cand_from_keys(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals, int64_t n,
                                 int bits_t, int bits_w, CandD* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const uint64_t k = keys[i];
    CandD x;
    x.mwidth = (int32_t)(k & ((1ull << bits_w) - 1));
    x.mt = (int32_t)((k >> bits_w) & ((1ull << bits_t) - 1));
    x.mdt = (int32_t)(k >> (bits_t + bits_w));
    x.msnr = __uint_as_float(vals[i]);
    out[i] = x;
}

__device__ __forceinline__ int32_t member_of(const CandD& c, int m) { return m == 0 ? c.mt : (m == 1 ? c.mdt : c.mwidth); }

// keys[i] = group[i] << 32 | ord(member(c[perm[i]])), vals[i] = perm[i]
__global__ void k_cluster_keys(const CandD* __restrict__ c, const uint32_t* __restrict__ perm,
                               const uint32_t* __restrict__ group, int64_t n, int member,
                               uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const uint32_t p = perm[i];
    const uint64_t g = group ? (uint64_t)group[i] : 0ull;
    keys[i] = (g << 32) | ord_i32(member_of(c[p], member));
    vals[i] = p;
}

// flag[i] = 1 where a new group starts: first item, another parent group, or
// a member gap larger than the bin (Clusterization.cu:854-893,930-998).
__global__ void k_cluster_flags(const uint64_t* __restrict__ keys, int64_t n, const int32_t* __restrict__ bins,
                                int level, uint32_t* __restrict__ flags) {
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    uint32_t f = 1u;
    if (i > 0) {
        const uint64_t a = keys[i - 1], b = keys[i];
        if ((a >> 32) == (b >> 32)) {
            const int64_t d = (int64_t)from_ord((uint32_t)b) - (int64_t)from_ord((uint32_t)a);
            f = d > (int64_t)bins[level] ? 1u : 0u;
        }
    }
    flags[i] = f;
}

// gid = inclusive scan of flags (1-based group id); best[g] = max of
// (ord(msnr) << 32 | ~index) -> max S/N, ties to the lowest index;
// start[g] = first sorted position of group g.
__global__ void k_cluster_delegates(const CandD* __restrict__ c, const uint32_t* __restrict__ perm,
                                    const uint32_t* __restrict__ gid, const uint32_t* __restrict__ flags, int64_t n,
                                    unsigned long long* __restrict__ best, uint32_t* __restrict__ start) {
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const uint32_t g = gid[i] - 1u;
    const uint32_t p = perm[i];
    const unsigned long long key = ((unsigned long long)ord_f32(c[p].msnr) << 32) | (unsigned long long)(0xffffffffu - p);
    atomicMax(&best[g], key);
    if (flags[i]) start[g] = (uint32_t)i;
}

__global__ void k_delegate_keys(const CandD* __restrict__ c, const unsigned long long* __restrict__ best,
                                int64_t ng, uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const int64_t g = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (g >= ng) return;
    const uint32_t d = 0xffffffffu - (uint32_t)(best[g] & 0xffffffffull);
    keys[g] = ((uint64_t)ord_i32(c[d].mt) << 32) | (uint64_t)g;
    vals[g] = (uint32_t)g;
}

__global__ void k_delegate_out(const unsigned long long* __restrict__ best, const uint32_t* __restrict__ start,
                               const uint32_t* __restrict__ order, int64_t ng, int64_t n,
                               int32_t* __restrict__ delegates, int64_t* __restrict__ sizes) {
    const int64_t j = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (j >= ng) return;
    const uint32_t g = order[j];
    delegates[j] = (int32_t)(0xffffffffu - (uint32_t)(best[g] & 0xffffffffull));
    const int64_t end = (g + 1 < (uint64_t)ng) ? (int64_t)start[g + 1] : n;
    sizes[j] = end - (int64_t)start[g];
}

__global__ void k_iota(uint32_t* v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i < n) v[i] = (uint32_t)i;
}

__global__ void k_fill_u64(unsigned long long* v, int64_t n, unsigned long long val) {
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i < n) v[i] = val;
}

inline uint32_t grid1(int64_t n) { return (uint32_t)ceil_div(n > 0 ? n : 1, kThreads); }

// select_delegates (select_delegates.cu:16-51,62-125) over caller-provided
// groups: group g is positions [begin[g], begin[g+1]) (the last ends at
// nidx) of the index array; its delegate is the candidate index there with
// the largest msnr, the first such position on ties -- the rule of
// thrust::max_element with a less-than comparator, which also keeps a NaN
// found at the group's first position.  One wave per group.
__global__ void __launch_bounds__(64) k_select_delegates(const CandD* __restrict__ c, const int32_t* __restrict__ idx,
                                                         int64_t nidx, const int32_t* __restrict__ begin,
                                                         int64_t ngroups, int32_t* __restrict__ out) {
    const int64_t g = blockIdx.x;
    const int64_t b = begin[g];
    const int64_t e = g + 1 < ngroups ? (int64_t)begin[g + 1] : nidx;
    if (b >= e) {
        if (threadIdx.x == 0) out[g] = -1;
        return;
    }
    const float first = c[idx[b]].msnr;
    if (!(first == first)) {
        if (threadIdx.x == 0) out[g] = idx[b];
        return;
    }
    float bv = first;
    int64_t bp = b;
    for (int64_t p = b + threadIdx.x; p < e; p += 64) {
        const float v = c[idx[p]].msnr;
        if (v > bv || (v == bv && p < bp)) {
            bv = v;
            bp = p;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        const float ov = __shfl_down(bv, off);
        const int64_t op = __shfl_down(bp, off);
        if (ov > bv || (ov == bv && op < bp)) {
            bv = ov;
            bp = op;
        }
    }
    if (threadIdx.x == 0) out[g] = idx[bp];
}

}  // namespace

// ---------------------------------------------------------------------------
int64_t scan_scratch_entries(int64_t n) {
    int64_t tot = 0;
    for (int64_t m = ceil_div(n, kScanBlock); m > 1; m = ceil_div(m, kScanBlock)) tot += m;
    return tot + 1;
}

template <typename T>
static void scan_inclusive(T* data, int64_t n, T* scratch, hipStream_t st) {
    if (n <= 0) return;
    const int64_t nb = ceil_div(n, kScanBlock);
    if (nb > 0x7fffffffLL) throw Error(SPS_ERR_INVALID_ARGUMENT, "scan too long");
    if (nb == 1) {
        hipLaunchKernelGGL(k_scan_local<T>, dim3(1), dim3(kThreads), 0, st, data, n, (T*)nullptr);
        SPS_HIP(hipGetLastError());
        return;
    }
    T* sums = scratch;
    hipLaunchKernelGGL(k_scan_local<T>, dim3((uint32_t)nb), dim3(kThreads), 0, st, data, n, sums);
    scan_inclusive<T>(sums, nb, scratch + nb, st);
    hipLaunchKernelGGL(k_scan_add<T>, dim3((uint32_t)nb), dim3(kThreads), 0, st, data, n, (const T*)sums);
    SPS_HIP(hipGetLastError());
}

void scan_inclusive_u32(uint32_t* data, int64_t n, uint32_t* scratch, hipStream_t st) {
    scan_inclusive<uint32_t>(data, n, scratch, st);
}

void scan_inclusive_f32(float* data, int64_t n, float* scratch, hipStream_t st) {
    scan_inclusive<float>(data, n, scratch, st);
}

int64_t radix_hist_entries(int64_t n) {
    const int64_t nb = ceil_div(n > 0 ? n : 1, kSortBlock);
    const int64_t h = nb * kRadix;
    return 2 * h + scan_scratch_entries(h);
}

void radix_sort_pairs(uint64_t* keys, uint32_t* vals, uint64_t* keys_tmp, uint32_t* vals_tmp, int64_t n, int nbits,
                      uint32_t* d_hist, int64_t hist_cap, hipStream_t st) {
    if (n <= 1 || nbits <= 0) return;
    if (n > 0xffffffffLL) throw Error(SPS_ERR_INVALID_ARGUMENT, "too many items to sort");
    const int64_t nb = ceil_div(n, kSortBlock);
    const int64_t h = nb * kRadix;
    if (radix_hist_entries(n) > hist_cap) throw Error(SPS_ERR_INVALID_ARGUMENT, "radix scratch too small");
    uint32_t* hist = d_hist;
    uint32_t* offs = d_hist + h;
    uint32_t* scan_scratch = d_hist + 2 * h;
    const int passes = (nbits + kRadixBits - 1) / kRadixBits;
    uint64_t* ki = keys;
    uint32_t* vi = vals;
    uint64_t* ko = keys_tmp;
    uint32_t* vo = vals_tmp;
    for (int p = 0; p < passes; ++p) {
        const int shift = p * kRadixBits;
        hipLaunchKernelGGL(k_radix_hist, dim3((uint32_t)nb), dim3(kThreads), 0, st, ki, n, shift, hist, nb);
        scan_inclusive_u32(hist, h, scan_scratch, st);
        hipLaunchKernelGGL(k_excl_from_incl, dim3(grid1(h)), dim3(kThreads), 0, st, hist, offs, h);
        hipLaunchKernelGGL(k_radix_scatter, dim3((uint32_t)nb), dim3(kThreads), 0, st, ki, vi, n, shift,
                           (const uint32_t*)offs, nb, ko, vo);
        SPS_HIP(hipGetLastError());
        std::swap(ki, ko);
        std::swap(vi, vo);
    }
    if (ki != keys) {
        SPS_HIP(hipMemcpyAsync(keys, ki, sizeof(uint64_t) * (size_t)n, hipMemcpyDeviceToDevice, st));
        SPS_HIP(hipMemcpyAsync(vals, vi, sizeof(uint32_t) * (size_t)n, hipMemcpyDeviceToDevice, st));
    }
}

void launch_cand_keys(const CandD* c, int64_t n, int bits_t, int bits_w, uint64_t* keys, uint32_t* vals,
                      hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_cand_keys, dim3(grid1(n)), dim3(kThreads), 0, st, c, n, bits_t, bits_w, keys, vals);
    SPS_HIP(hipGetLastError());
}

void launch_cand_from_keys(const uint64_t* keys, const uint32_t* vals, int64_t n, int bits_t, int bits_w, CandD* out,
                           hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_cand_from_keys, dim3(grid1(n)), dim3(kThreads), 0, st, keys, vals, n, bits_t, bits_w, out);
    SPS_HIP(hipGetLastError());
}

void launch_cluster_keys(const CandD* c, const uint32_t* perm, const uint32_t* group, int64_t n, int member,
                         uint64_t* keys, uint32_t* vals, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_cluster_keys, dim3(grid1(n)), dim3(kThreads), 0, st, c, perm, group, n, member, keys, vals);
    SPS_HIP(hipGetLastError());
}

void launch_cluster_flags(const uint64_t* keys, int64_t n, const int32_t* d_bins, int level, uint32_t* flags,
                          hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_cluster_flags, dim3(grid1(n)), dim3(kThreads), 0, st, keys, n, d_bins, level, flags);
    SPS_HIP(hipGetLastError());
}

void launch_cluster_delegates(const CandD* c, const uint32_t* perm, const uint32_t* gid, const uint32_t* flags,
                              int64_t n, unsigned long long* best, uint32_t* start, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_cluster_delegates, dim3(grid1(n)), dim3(kThreads), 0, st, c, perm, gid, flags, n, best,
                       start);
    SPS_HIP(hipGetLastError());
}

void launch_delegate_keys(const CandD* c, const unsigned long long* best, const uint32_t* start, int64_t ng,
                          int64_t n, uint64_t* keys, uint32_t* vals, hipStream_t st) {
    (void)start;
    (void)n;
    if (ng <= 0) return;
    hipLaunchKernelGGL(k_delegate_keys, dim3(grid1(ng)), dim3(kThreads), 0, st, c, best, ng, keys, vals);
    SPS_HIP(hipGetLastError());
}

void launch_delegate_out(const unsigned long long* best, const uint32_t* start, const uint32_t* order, int64_t ng,
                         int64_t n, int32_t* delegates, int64_t* sizes, hipStream_t st) {
    if (ng <= 0) return;
    hipLaunchKernelGGL(k_delegate_out, dim3(grid1(ng)), dim3(kThreads), 0, st, best, start, order, ng, n, delegates,
                       sizes);
    SPS_HIP(hipGetLastError());
}

void launch_iota_u32(uint32_t* v, int64_t n, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_iota, dim3(grid1(n)), dim3(kThreads), 0, st, v, n);
    SPS_HIP(hipGetLastError());
}

void launch_fill_u64(unsigned long long* v, int64_t n, unsigned long long val, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_u64, dim3(grid1(n)), dim3(kThreads), 0, st, v, n, val);
    SPS_HIP(hipGetLastError());
}

void launch_select_delegates(const CandD* c, const int32_t* idx, int64_t nidx, const int32_t* begin, int64_t ngroups,
                             int32_t* out, hipStream_t st) {
    if (ngroups <= 0) return;
    if (ngroups > 0x7fffffffLL) throw Error(SPS_ERR_INVALID_ARGUMENT, "too many groups");
    hipLaunchKernelGGL(k_select_delegates, dim3((uint32_t)ngroups), dim3(64), 0, st, c, idx, nidx, begin, ngroups, out);
    SPS_HIP(hipGetLastError());
}

}  // namespace sps
