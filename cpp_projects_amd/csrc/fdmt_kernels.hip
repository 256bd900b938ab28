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
#include <type_traits>

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
// SNR: the output is divided by sqrtf(n + 1e-8) (fdmt_normalization,
// Chunk_gpu.cu:841-852) with n from the mask transform for t < nrm_len and the
// kept-channel count beyond it.  AUX only renames the instantiation used for
// the mask transform so profiles keep it apart from the data transform.
struct SnrEpi {
    const float* nrm;     // rows x nrm_len
    int64_t nrm_len;
    const double* stats;  // stats[4] = kept channels
};

template <bool VEC, bool SNR, bool AUX>
__global__ void __launch_bounds__(kThreads) k_merge(const float* __restrict__ in, int64_t in_pitch,
                                                    float* __restrict__ out, int64_t out_pitch,
                                                    const RowDescD* __restrict__ rows, int64_t nrows,
                                                    int64_t nsamps, int64_t out_rows_limit, SnrEpi epi) {
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
    float dk = 1.0f, rdk = 1.0f;
    const float* nrow = nullptr;
    if (SNR) {
        dk = sqrtf((float)((double)(float)epi.stats[4] + 1.0E-8));
        rdk = 1.0f / dk;
        nrow = epi.nrm + d.out_row * epi.nrm_len;
    }
    // beyond the mask transform's extent the divisor is the same for every
    // sample, so its reciprocal is shared (correctly rounded, div_rn)
    auto div = [&](float v, int64_t t) -> float {
        if (!SNR) return v;
        if (t < epi.nrm_len) return v / sqrtf((float)((double)nrow[t] + 1.0E-8));
        return div_rn(v, dk, rdk);
    };
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
                if (SNR) {
                    a.x = div(a.x, t + 0);
                    a.y = div(a.y, t + 1);
                    a.z = div(a.z, t + 2);
                    a.w = div(a.w, t + 3);
                }
                *reinterpret_cast<float4a*>(o + t) = a;
            } else {
                for (int e = 0; e < 4; ++e) {
                    const int64_t tt = t + e;
                    if (tt < nsamps) {
                        float v = tail[tt];
                        if (has_head && tt >= off) v = v + head[tt - off];
                        o[tt] = div(v, tt);
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
                o[t] = div(v, t);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// fused passes: L consecutive iterations in one launch
// ---------------------------------------------------------------------------
// Every output row of iteration i+L-1 is the root of a binary tree in heap
// order: internal node k (1 <= k < 2^L) holds the head offset of its merge
// (or -1 when the row is a copy), its tail child is 2k and its head child
// 2k+1; the leaves 2^L..2^(L+1)-1 are rows of state i-1.  Evaluating
// v(k,t) = v(2k,t) + (t >= off_k ? v(2k+1, t-off_k) : 0) performs exactly the
// fp32 additions of the L separate iterations, in the same order, so the
// intermediate states are never written.
template <int L, int K>
struct TreeEval {
    // scalar value of node K at time t (t >= 0)
    __device__ __forceinline__ static float one(const float* __restrict__ in, int64_t pitch, const TreeDesc<L>& d,
                                                int64_t t) {
        if constexpr (K >= (1 << L)) {
            return in[d.leaf[K - (1 << L)] * pitch + t];
        } else {
            float v = TreeEval<L, 2 * K>::one(in, pitch, d, t);
            const int64_t off = d.off[K];
            if (off >= 0 && t >= off) v = v + TreeEval<L, 2 * K + 1>::one(in, pitch, d, t - off);
            return v;
        }
    }
    // four consecutive times t..t+3 (t >= 0); leaf loads are 4-byte aligned
    // 16-byte loads, except the root tail chain which is 16-byte aligned
    __device__ __forceinline__ static float4a four(const float* __restrict__ in, int64_t pitch, const TreeDesc<L>& d,
                                                   int64_t t) {
        if constexpr (K >= (1 << L)) {
            const float4u v = *reinterpret_cast<const float4u*>(in + d.leaf[K - (1 << L)] * pitch + t);
            float4a r;
            r.x = v.x;
            r.y = v.y;
            r.z = v.z;
            r.w = v.w;
            return r;
        } else {
            float4a v = TreeEval<L, 2 * K>::four(in, pitch, d, t);
            const int64_t off = d.off[K];
            if (off >= 0) {
                if (t >= off) {
                    const float4a h = TreeEval<L, 2 * K + 1>::four(in, pitch, d, t - off);
                    v.x = v.x + h.x;
                    v.y = v.y + h.y;
                    v.z = v.z + h.z;
                    v.w = v.w + h.w;
                } else if (t + 3 >= off) {
                    if (t + 0 >= off) v.x = v.x + TreeEval<L, 2 * K + 1>::one(in, pitch, d, t + 0 - off);
                    if (t + 1 >= off) v.y = v.y + TreeEval<L, 2 * K + 1>::one(in, pitch, d, t + 1 - off);
                    if (t + 2 >= off) v.z = v.z + TreeEval<L, 2 * K + 1>::one(in, pitch, d, t + 2 - off);
                    if (t + 3 >= off) v.w = v.w + TreeEval<L, 2 * K + 1>::one(in, pitch, d, t + 3 - off);
                }
            }
            return v;
        }
    }
};

template <int L, bool VEC, bool SNR>
__global__ void __launch_bounds__(kThreads) k_merge_fused(const float* __restrict__ in, int64_t in_pitch,
                                                          float* __restrict__ out, int64_t out_pitch,
                                                          const TreeDesc<L>* __restrict__ rows, int64_t nrows,
                                                          int64_t nsamps, int64_t out_rows_limit, SnrEpi epi) {
    const uint32_t logical = xcd_remap(blockIdx.x, gridDim.x);
    const int64_t ri = (int64_t)(logical % (uint64_t)nrows);
    const int64_t tile = (int64_t)(logical / (uint64_t)nrows);
    const TreeDesc<L> d = rows[ri];
    if (d.out_row >= out_rows_limit) return;
    float* __restrict__ o = out + d.out_row * out_pitch;
    const int64_t base = tile * kMergeTile;
    float dk = 1.0f, rdk = 1.0f;
    const float* nrow = nullptr;
    if (SNR) {
        dk = sqrtf((float)((double)(float)epi.stats[4] + 1.0E-8));
        rdk = 1.0f / dk;
        nrow = epi.nrm + d.out_row * epi.nrm_len;
    }
    // beyond the mask transform's extent the divisor is the same for every
    // sample, so its reciprocal is shared (correctly rounded, div_rn)
    auto div = [&](float v, int64_t t) -> float {
        if (!SNR) return v;
        if (t < epi.nrm_len) return v / sqrtf((float)((double)nrow[t] + 1.0E-8));
        return div_rn(v, dk, rdk);
    };
    if (VEC) {
#pragma unroll
        for (int j = 0; j < kMergeVecPerThread; ++j) {
            const int64_t t = base + (int64_t)j * (kThreads * 4) + (int64_t)threadIdx.x * 4;
            if (t + 4 <= nsamps) {
                float4a a = TreeEval<L, 1>::four(in, in_pitch, d, t);
                if (SNR) {
                    a.x = div(a.x, t + 0);
                    a.y = div(a.y, t + 1);
                    a.z = div(a.z, t + 2);
                    a.w = div(a.w, t + 3);
                }
                *reinterpret_cast<float4a*>(o + t) = a;
            } else {
                for (int e = 0; e < 4; ++e)
                    if (t + e < nsamps) o[t + e] = div(TreeEval<L, 1>::one(in, in_pitch, d, t + e), t + e);
            }
        }
    } else {
        for (int j = 0; j < kMergeTile / kThreads; ++j) {
            const int64_t t = base + (int64_t)j * kThreads + threadIdx.x;
            if (t < nsamps) o[t] = div(TreeEval<L, 1>::one(in, in_pitch, d, t), t);
        }
    }
}

// ---------------------------------------------------------------------------
// LDS-staged passes: one block = (group of consecutive output rows, 512-sample
// tile).  Every distinct leaf row of the group is read from HBM once per tile
// (aligned float4 units, each wave staging its own leaves with up to 12 loads
// in flight per lane); the trees of all rows are then evaluated from LDS with
// the same fp32 additions, in the same order, as the separate iterations.
// ---------------------------------------------------------------------------
template <int L, int K>
struct LdsTree {
    __device__ __forceinline__ static float eval(const float* __restrict__ lds, const GRowDesc<L>& d, int64_t t,
                                                 int64_t T) {
        if constexpr (K >= (1 << L)) {
            return lds[d.lbase[K - (1 << L)] + (int)(t - T)];
        } else {
            float v = LdsTree<L, 2 * K>::eval(lds, d, t, T);
            const int32_t off = d.off[K];
            if (off >= 0 && t >= off) v = v + LdsTree<L, 2 * K + 1>::eval(lds, d, t - off, T);
            return v;
        }
    }
};

// the same tree when every merge adds its head (t >= maxshift): u = t - T.
// Branch-free, so all 2^L leaf reads can be in flight at once: a missing
// head (a copied row, off < 0) is still read, at offset 0 from leaves whose
// base is 0, and dropped by the select; the kept additions are those of the
// separate iterations.  FULL: the row has no missing head (GRowDesc::full),
// so no select is needed.
template <int L, int K, bool FULL = false>
struct LdsTreeFast {
    __device__ __forceinline__ static float eval(const float* __restrict__ lds, const GRowDesc<L>& d, int u) {
        if constexpr (K >= (1 << L)) {
            return lds[d.lbase[K - (1 << L)] + u];
        } else {
            const int32_t off = d.off[K];
            const float a = LdsTreeFast<L, 2 * K, FULL>::eval(lds, d, u);
            if constexpr (FULL) {
                return a + LdsTreeFast<L, 2 * K + 1, FULL>::eval(lds, d, u - off);
            } else {
                const float b = LdsTreeFast<L, 2 * K + 1, FULL>::eval(lds, d, u - (off >= 0 ? off : 0));
                return off >= 0 ? a + b : a;
            }
        }
    }
};

// evaluates the trees of every row of group g for times [T, T + TILE): wave w
// takes rows w, w + 4, ... over the whole tile (TILE / 64 samples per lane), so
// the per-row scalar work (descriptor, output address, path test) is done
// once per row rather than once per wave
// FAST: the caller guarantees that the tile is whole and past every row's
// shifts (T >= maxshift) and, with SNR, past the mask transform's extent, so
// only the branch-free path is compiled (the general one's registers made the
// hot loop spill).
// W0, NW: the rows go to waves W0 .. W0 + NW - 1 of the block.
template <int L, bool SNR, bool FAST = false, int TILE = kGroupTile, int W0 = 0, int NW = kThreads / 64>
__device__ __forceinline__ void group_trees(const float* __restrict__ lds, const GroupDescD& g,
                                            const GRowDesc<L>* __restrict__ rows, float* __restrict__ out,
                                            int64_t out_pitch, int64_t T, int64_t nsamps, int64_t out_rows_limit,
                                            const SnrEpi& epi) {
    constexpr int kWaves = NW;
    const int lane = (int)threadIdx.x & 63;
    const int wid = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6) - W0;
    float dk = 1.0f, rdk = 1.0f;
    if (SNR) {
        dk = sqrtf((float)((double)(float)epi.stats[4] + 1.0E-8));
        rdk = 1.0f / dk;
    }
    for (int i = wid; i < g.nrows; i += kWaves) {  // wave-uniform
        const GRowDesc<L> d = rows[g.row0 + i];
        if (d.out_row >= out_rows_limit) continue;
        float* __restrict__ o = out + d.out_row * out_pitch;
        const float* nrow = SNR ? epi.nrm + d.out_row * epi.nrm_len : nullptr;
        if (FAST || (T >= d.maxshift && T + TILE <= nsamps)) {
            // the whole tile is past every shift: no per-sample tests.  The
            // outputs are stored non-temporally: a state is far larger than
            // the caches and is read back only by the next pass (measured:
            // 0.2 ms less per search step at 4096 x 2^18)
            auto fast = [&](auto full_tag) {
                constexpr bool FULL = decltype(full_tag)::value;
                if (!FAST && SNR && T < epi.nrm_len) {
                    // within the mask transform's extent: per-sample divisor
                    // there, all of the row-tile's divisors loaded first
                    float nv[TILE / 64];
#pragma unroll
                    for (int j = 0; j < TILE / 64; ++j) {
                        const int64_t t = T + j * 64 + lane;
                        nv[j] = (t < epi.nrm_len) ? nrow[t] : 0.0f;
                    }
#pragma unroll 4
                    for (int j = 0; j < TILE / 64; ++j) {
                        const int u = j * 64 + lane;
                        const int64_t t = T + u;
                        const float v = LdsTreeFast<L, 1, FULL>::eval(lds, d, u);
                        o[t] = (t < epi.nrm_len) ? v / sqrtf((float)((double)nv[j] + 1.0E-8)) : div_rn(v, dk, rdk);
                    }
                } else if constexpr (SNR) {
                    // div_rn(v, dk, rdk) four samples at a time: dk = sqrtf(kept +
                    // 1e-8) lies in [1e-4, 256], so only the dividends need the
                    // range test, and the IEEE division runs in one rare branch
#pragma unroll
                    for (int j0 = 0; j0 < TILE / 64; j0 += 4) {
                        float v[4], qv[4];
                        bool slow = false;
#pragma unroll
                        for (int jj = 0; jj < 4; ++jj) v[jj] = LdsTreeFast<L, 1, FULL>::eval(lds, d, (j0 + jj) * 64 + lane);
#pragma unroll
                        for (int jj = 0; jj < 4; ++jj) {
                            const float ax = fabsf(v[jj]);
                            slow = slow || !(ax >= 0x1p-60f && ax <= 0x1p60f);
                            const float q0 = v[jj] * rdk;
                            const float e = fmaf(-q0, dk, v[jj]);
                            qv[jj] = fmaf(e, rdk, q0);
                        }
                        if (slow) {
#pragma unroll
                            for (int jj = 0; jj < 4; ++jj) {
                                const float ax = fabsf(v[jj]);
                                if (!(ax >= 0x1p-60f && ax <= 0x1p60f)) qv[jj] = v[jj] / dk;
                            }
                        }
#pragma unroll
                        for (int jj = 0; jj < 4; ++jj) __builtin_nontemporal_store(qv[jj], &o[T + (j0 + jj) * 64 + lane]);
                    }
                } else {
#pragma unroll 4
                    for (int j = 0; j < TILE / 64; ++j) {
                        const int u = j * 64 + lane;
                        __builtin_nontemporal_store(LdsTreeFast<L, 1, FULL>::eval(lds, d, u), &o[T + u]);
                    }
                }
            };
            if (d.full)
                fast(std::true_type{});
            else
                fast(std::false_type{});
            continue;
        }
        if constexpr (!FAST) {
            float nv[TILE / 64];
            if (SNR) {
#pragma unroll
                for (int j = 0; j < TILE / 64; ++j) {
                    const int64_t t = T + j * 64 + lane;
                    nv[j] = (t < epi.nrm_len && t < nsamps) ? nrow[t] : 0.0f;
                }
            }
#pragma unroll 4
            for (int j = 0; j < TILE / 64; ++j) {
                const int64_t t = T + j * 64 + lane;
                if (t < nsamps) {
                    float v = LdsTree<L, 1>::eval(lds, d, t, T);
                    if (SNR)
                        v = (t < epi.nrm_len) ? v / sqrtf((float)((double)nv[j] + 1.0E-8)) : div_rn(v, dk, rdk);
                    o[t] = v;
                }
            }
        }
    }
}

// One block = one group x tiles_per_block consecutive tiles of [tile_lo,
// tile_hi) (FAST: tiles where group_trees may take its fast path for every
// row, see there).  Every float4
// unit of the group's LDS windows is a fixed (leaf row, offset) pair, so each
// thread works out its unit slots once; per tile it issues the loads of the
// next tile into registers, evaluates the current tile from one LDS buffer,
// then commits the registers to the other buffer (one barrier per tile).
// ONEBUF: one LDS buffer, committed after a second barrier once the trees
// have read it (half the LDS, so wide windows keep more blocks resident).
constexpr int kStageSlots = kGroupMaxUnits / kThreads;

template <int L, bool SNR, bool FAST, bool ONEBUF>
__global__ void __launch_bounds__(kThreads) k_merge_group(const float* __restrict__ in, int64_t in_pitch,
                                                          float* __restrict__ out, int64_t out_pitch,
                                                          const GroupDescD* __restrict__ groups, int64_t ngroups,
                                                          const LeafDescD* __restrict__ leaves,
                                                          const GRowDesc<L>* __restrict__ rows,
                                                          const uint8_t* __restrict__ unit_leaf, int tiles_per_block,
                                                          int lds_stride, int64_t nsamps, int64_t out_rows_limit,
                                                          SnrEpi epi, int64_t tile_lo, int64_t tile_hi) {
    extern __shared__ __attribute__((aligned(16))) float lds[];
    const uint32_t logical = xcd_remap(blockIdx.x, gridDim.x);
    const int64_t gi = (int64_t)(logical % (uint64_t)ngroups);
    const int64_t chunk = (int64_t)(logical / (uint64_t)ngroups);
    const GroupDescD g = groups[gi];
    const int64_t k0 = tile_lo + chunk * tiles_per_block;
    const int64_t k1 = min(tile_hi, k0 + tiles_per_block);
    if (k0 >= k1) return;  // block-uniform
    int64_t src[kStageSlots];  // element offset of the slot's unit at T = 0
    int32_t tk[kStageSlots];   // its time at T = 0; INT32_MIN for an unused slot
#pragma unroll
    for (int j = 0; j < kStageSlots; ++j) {
        const int f = (int)threadIdx.x + j * kThreads;
        src[j] = 0;
        tk[j] = INT32_MIN;
        if (f < g.nunits) {
            const LeafDescD lf = leaves[g.leaf0 + unit_leaf[g.unit0 + f]];
            tk[j] = 4 * (f - lf.lds_off / 4) - lf.smax4;
            src[j] = lf.row * in_pitch + tk[j];
        }
    }
    float4a buf[kStageSlots];
    auto issue = [&](int64_t T) {
#pragma unroll
        for (int j = 0; j < kStageSlots; ++j) {
            const int64_t t = T + tk[j];
            float4a v = {0.0f, 0.0f, 0.0f, 0.0f};
            if (tk[j] != INT32_MIN && t >= 0 && t + 4 <= in_pitch) v = *reinterpret_cast<const float4a*>(in + src[j] + T);
            buf[j] = v;
        }
    };
    auto commit = [&](float* dst) {
#pragma unroll
        for (int j = 0; j < kStageSlots; ++j) {
            const int f = (int)threadIdx.x + j * kThreads;
            if (f < g.nunits) *reinterpret_cast<float4a*>(dst + 4 * f) = buf[j];
        }
    };
    issue(k0 * kGroupTile);
    commit(lds);
    __syncthreads();
    for (int64_t k = k0; k < k1; ++k) {
        float* cur = ONEBUF ? lds : lds + ((k - k0) & 1) * lds_stride;
        float* nxt = ONEBUF ? lds : lds + ((k - k0 + 1) & 1) * lds_stride;
        const bool more = k + 1 < k1;
        if (more) issue((k + 1) * kGroupTile);
        group_trees<L, SNR, FAST>(cur, g, rows, out, out_pitch, k * kGroupTile, nsamps, out_rows_limit, epi);
        if (ONEBUF && more) __syncthreads();  // block-uniform: every read of cur is done
        if (more) commit(nxt);
        __syncthreads();
    }
}

// Producer/consumer form of the fast launch (tiles whose every unit lies in
// [0, in_pitch): T is a multiple of 4 at or past every leaf's rounded shift,
// and the window ends at most 3 samples past T + 512 <= nsamps, inside the
// 4-sample padded row).  Wave 0 only loads: per tile it copies the group's
// units of the next tile straight into the other LDS buffer with
// global_load_lds_dwordx4 and waits for its own loads alone; waves 1 .. 7
// only evaluate trees and store.  In k_merge_group every wave waits for
// the whole vector-memory counter before committing the next tile, and on
// this GPU that counter also holds the tile's output stores, so each tile
// waits for its stores to drain; here the consumers never wait on memory.
constexpr int kPCThreads = 512;
constexpr int kPCUnits = kGroupMaxUnits / 64;  // units per producer lane

template <int L, bool SNR>
__global__ void __launch_bounds__(kPCThreads) k_merge_group_pc(const float* __restrict__ in, int64_t in_pitch,
                                                               float* __restrict__ out, int64_t out_pitch,
                                                               const GroupDescD* __restrict__ groups,
                                                               int64_t ngroups, const LeafDescD* __restrict__ leaves,
                                                               const GRowDesc<L>* __restrict__ rows,
                                                               const uint8_t* __restrict__ unit_leaf,
                                                               int tiles_per_block, int lds_stride, int64_t nsamps,
                                                               int64_t out_rows_limit, SnrEpi epi, int64_t tile_lo,
                                                               int64_t tile_hi) {
    extern __shared__ __attribute__((aligned(16))) float lds[];
    const uint32_t logical = xcd_remap(blockIdx.x, gridDim.x);
    const int64_t gi = (int64_t)(logical % (uint64_t)ngroups);
    const int64_t chunk = (int64_t)(logical / (uint64_t)ngroups);
    const GroupDescD g = groups[gi];
    const int64_t k0 = tile_lo + chunk * tiles_per_block;
    const int64_t k1 = min(tile_hi, k0 + tiles_per_block);
    if (k0 >= k1) return;  // block-uniform
    // the group's leaves (at most 64) in LDS for the producer's set-up
    __shared__ LeafDescD lfs[64];
    if ((int)threadIdx.x < g.nleaves) lfs[threadIdx.x] = leaves[g.leaf0 + threadIdx.x];
    __syncthreads();
    const int wid = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    if (wid == 0) {
        const int lane = (int)threadIdx.x & 63;
        // lane l, slot i: unit 64 i + l, its element offset at T = 0 (the
        // table reads are unconditional, clamped to the group's units, so
        // they are all in flight together)
        uint32_t li[kPCUnits];
#pragma unroll
        for (int i = 0; i < kPCUnits; ++i) li[i] = unit_leaf[g.unit0 + min(lane + 64 * i, g.nunits - 1)];
        int64_t src[kPCUnits];
#pragma unroll
        for (int i = 0; i < kPCUnits; ++i) {
            const LeafDescD& lf = lfs[li[i]];
            src[i] = lf.row * in_pitch + (4 * (lane + 64 * i) - lf.lds_off) - lf.smax4;
        }
        // slot i of the wave fills LDS bytes [1024 i, 1024 (i + 1)) of the
        // buffer: lane l's 16 bytes at M0 + 16 l.  Issued as inline asm, so
        // the compiler counts none of it; the wave waits itself.
        auto issue = [&](int64_t T, float* dst) {
#pragma unroll
            for (int i = 0; i < kPCUnits; ++i) {
                if (64 * i >= g.nunits) break;  // wave-uniform
                const uint32_t m0 = __builtin_amdgcn_readfirstlane(
                    (uint32_t)(uintptr_t)(__attribute__((address_space(3))) float*)(dst + 256 * i));
                const float* gp = in + src[i] + T;
                uint32_t saved;
                if (lane + 64 * i < g.nunits)
                    asm volatile("s_mov_b32 %0, m0\n\ts_mov_b32 m0, %1\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %2, off\n\t"
                                 "s_mov_b32 m0, %0"
                                 : "=&s"(saved)
                                 : "s"(m0), "v"(gp)
                                 : "memory");
            }
        };
        issue(k0 * kGroupTile, lds);
        asm volatile("s_waitcnt vmcnt(0)" : : : "memory");
        __syncthreads();
        for (int64_t k = k0; k < k1; ++k) {
            if (k + 1 < k1) {
                issue((k + 1) * kGroupTile, lds + ((k - k0 + 1) & 1) * lds_stride);
                asm volatile("s_waitcnt vmcnt(0)" : : : "memory");
            }
            __syncthreads();  // the next tile has landed; the consumers are done with this one
        }
    } else {
        __syncthreads();
        for (int64_t k = k0; k < k1; ++k) {
            const float* cur = lds + ((k - k0) & 1) * lds_stride;
            group_trees<L, SNR, true, kGroupTile, 1, kPCThreads / 64 - 1>(cur, g, rows, out, out_pitch,
                                                                          k * kGroupTile, nsamps, out_rows_limit, epi);
            __syncthreads();
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
                                                   float* __restrict__ state, int64_t pitch,
                                                   const uint8_t* __restrict__ live) {
    const int64_t c = blockIdx.y;
    const int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (t >= nsamps) return;
    const uint8_t* lv = live ? live + chan_row_off[c] : nullptr;
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
    if (!lv || lv[0]) row[t] = prev;
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
        if (!lv || lv[k]) row[(int64_t)k * pitch + t] = v;
        prev = v;
    }
}

// LDS-staged init: one block per (channel, 4096-sample tile).  The tile plus
// its history of dmax samples is read once (16-byte loads when aligned),
// normalised once into LDS, and every thread then produces 16 lane-strided
// samples of each of the channel's rows with the same operation order as
// k_init (rows that no output depends on are computed but not stored).
constexpr int kInitTile = 4096;

template <int IN_KIND>
__device__ __forceinline__ void load_span(const void* __restrict__ in, int64_t row_base, int64_t lo, int n,
                                          float* __restrict__ dst, bool aligned16) {
    // dst[i] = raw(lo + i), i in [0, n); lo may be negative (zeros)
    if (IN_KIND == 2) {
        for (int i = threadIdx.x; i < n; i += kThreads) dst[i] = 1.0f;
        return;
    }
    if (IN_KIND == 1 && aligned16) {
        const uint8_t* p = static_cast<const uint8_t*>(in) + row_base;
        const int64_t a0 = (lo >= 0 ? lo : 0) & ~(int64_t)15;
        const int64_t hi = lo + n;
        for (int64_t c = a0 + (int64_t)threadIdx.x * 16; c < hi; c += (int64_t)kThreads * 16) {
            const uint4 v = *reinterpret_cast<const uint4*>(p + c);
            const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                const int64_t idx = c + e;
                if (idx >= lo && idx < hi) dst[idx - lo] = (float)((w[e >> 2] >> (8 * (e & 3))) & 0xffu);
            }
        }
        for (int i = threadIdx.x; i < n && lo + i < 0; i += kThreads) dst[i] = 0.0f;
        return;
    }
    if (IN_KIND == 0 && aligned16) {
        const float* p = static_cast<const float*>(in) + row_base;
        const int64_t a0 = (lo >= 0 ? lo : 0) & ~(int64_t)3;
        const int64_t hi = lo + n;
        for (int64_t c = a0 + (int64_t)threadIdx.x * 4; c < hi; c += (int64_t)kThreads * 4) {
            const float4a v = *reinterpret_cast<const float4a*>(p + c);
            const float e4[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int64_t idx = c + e;
                if (idx >= lo && idx < hi) dst[idx - lo] = e4[e];
            }
        }
        for (int i = threadIdx.x; i < n && lo + i < 0; i += kThreads) dst[i] = 0.0f;
        return;
    }
    for (int i = threadIdx.x; i < n; i += kThreads) {
        const int64_t idx = lo + i;
        dst[i] = idx >= 0 ? load_in<IN_KIND>(in, row_base + idx) : 0.0f;
    }
}

template <int MODE, int IN_KIND, bool NORM>
__global__ void __launch_bounds__(kThreads) k_init_lds(const void* __restrict__ in, int64_t in_pitch, ChanNorm norm,
                                                       int64_t nsamps, const int32_t* __restrict__ grid,
                                                       const int64_t* __restrict__ grid_off,
                                                       const int64_t* __restrict__ chan_row_off,
                                                       float* __restrict__ state, int64_t pitch, int dmax_all,
                                                       int aligned_in, int aligned_out,
                                                       const uint8_t* __restrict__ live) {
    extern __shared__ __attribute__((aligned(16))) float xs[];
    const int64_t c = blockIdx.y;
    const int64_t T = (int64_t)blockIdx.x * kInitTile;
    const int64_t g0 = grid_off[c];
    const uint8_t* lv = live ? live + chan_row_off[c] : nullptr;
    const int K = (int)(grid_off[c + 1] - g0);
    const int32_t* g = grid + g0;
    const int dmax = g[K - 1];
    const int64_t n_t = min((int64_t)kInitTile, nsamps - T);
    const int n = (int)(n_t + dmax);
    // xs[i] holds sample T - dmax + i
    load_span<IN_KIND>(in, c * in_pitch, T - dmax, n, xs, aligned_in != 0);
    __syncthreads();
    if (NORM) {
        const bool keep = norm.keep[c] != 0;
        const float mu = norm.mu[c], sg = norm.sigma[c];
        const float rsg = 1.0f / sg;
        for (int i = threadIdx.x; i < n; i += kThreads) xs[i] = keep ? div_rn(xs[i] - mu, sg, rsg) : 0.0f;
        __syncthreads();
    }
    (void)dmax_all;
    (void)aligned_out;
    // rows outer, the thread's 16 samples inner (lane-consecutive, so the LDS
    // reads are conflict-free and each store instruction writes 256 bytes);
    // the running value of every sample stays in a register between rows
    constexpr int kPer = kInitTile / kThreads;
    float* rowbase = state + chan_row_off[c] * pitch;
    float prev[kPer];
    {
        const int64_t d0 = g[0];
        const bool keep_row = !lv || lv[0];
        const float y0 = (float)(d0 + 1), r0 = 1.0f / y0;
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const int lt = j * kThreads + threadIdx.x;
            const int64_t t = T + lt;
            const int xi = lt + dmax;  // LDS index of sample t
            float p = 0.0f;
            if (lt < n_t && t >= d0) {
                if (MODE == kInitCFdmtD) {
                    p = xs[xi];
                } else {
                    float sum = 0.0f;
                    for (int i = xi - (int)d0; i <= xi; ++i) sum += xs[i];
                    p = div_rn(sum, y0, r0);
                }
            }
            prev[j] = p;
            if (keep_row && lt < n_t) rowbase[t] = p;
        }
    }
    for (int k = 1; k < K; ++k) {
        const int64_t dc = g[k], dp = g[k - 1];
        const bool keep_row = !lv || lv[k];
        float* row = rowbase + (int64_t)k * pitch;
        const float yk = (float)dc + 1.0f, rk = 1.0f / yk;  // == (float)(dc + 1)
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const int lt = j * kThreads + threadIdx.x;
            const int64_t t = T + lt;
            const int xi = lt + dmax;
            float r = 0.0f;
            if (lt < n_t && t >= dc) {
                if (MODE == kInitMeanD) {
                    float sum = 0.0f;
                    for (int i = xi - (int)dc; i < xi - (int)dp; ++i) sum += xs[i];
                    r = div_rn(prev[j] * ((float)dp + 1.0f) + sum, yk, rk);
                } else if (MODE == kInitRestrictD) {
                    const float sum = 0.0f + xs[xi - (int)dc];
                    r = div_rn(prev[j] * (((float)dp + 1.0f) + sum), yk, rk);
                } else {
                    r = div_rn(prev[j] * (float)dc + xs[xi - (int)dc], yk, rk);
                }
            }
            prev[j] = r;
            if (keep_row && lt < n_t) row[t] = r;
        }
    }
}

// Mean-init fast path for grids whose init dts are 0, 1, ..., K-1 on every
// channel (K <= kInitConsecMax): row 0 is (0 + x[t]) / 1 = 0 + x[t] and row k
// adds the single sample x[t - k]: v_k = (v_{k-1} * k + (0 + x[t - k])) / (k + 1),
// the operations of k_init_lds for such a grid, in the same order.  One
// block = one channel x tiles_per_block consecutive 4096-sample tiles; a
// thread owns 4 chunks of 4 consecutive samples of a tile (chunk j * 256 +
// tid): one 4-byte (u8) or 16-byte (f32) load per chunk (none for ones), issued for the next
// tile before the current one is computed, normalised once in registers; the
// 4 samples before a chunk (x[t - k], k <= 4) come through one aligned
// 16-byte LDS read (for a tile's first chunk: the previous tile's last chunk,
// or the input before the block's first tile), and each row is written as
// float4, so every store instruction covers 1 KB (non-temporally: state 0 is
// read back only by the first merge pass).
constexpr int kInitConsecMax = 5;
constexpr int kInitTilesPerBlock = 8;

template <int IN_KIND, bool NORM>
__global__ void __launch_bounds__(kThreads) k_init_consec(const void* __restrict__ in, int64_t in_pitch, ChanNorm norm,
                                                          int64_t nsamps, const int64_t* __restrict__ grid_off,
                                                          const int64_t* __restrict__ chan_row_off,
                                                          float* __restrict__ state, int64_t pitch,
                                                          const uint8_t* __restrict__ live, int tiles_per_block) {
    static_assert(IN_KIND >= 0 && IN_KIND <= 2, "f32, u8 or ones input");
    __shared__ float4a xs[1 + kInitTile / 4];  // xs[0]: samples T-4 .. T-1; xs[1 + q]: chunk q
    const int64_t c = blockIdx.y;
    const int64_t ntiles = ceil_div(nsamps, (int64_t)kInitTile);
    const int64_t k0 = (int64_t)blockIdx.x * tiles_per_block;
    const int64_t k1 = min(k0 + (int64_t)tiles_per_block, ntiles);
    const int K = (int)(grid_off[c + 1] - grid_off[c]);
    // which rows are written, read once: a uint8_t load may alias the float
    // stores below, so reading the flags per store would reload them each time
    uint32_t wmask = 0u;
    for (int kk = 0; kk < K && kk < kInitConsecMax; ++kk)
        if (!live || live[chan_row_off[c] + kk]) wmask |= 1u << kk;
    float mu = 0.0f, sg = 1.0f, rsg = 1.0f;
    bool keep = true;
    if (NORM) {
        keep = norm.keep[c] != 0;
        mu = norm.mu[c];
        sg = norm.sigma[c];
        rsg = 1.0f / sg;
    }
    auto nrm = [&](float v) -> float { return NORM ? (keep ? div_rn(v - mu, sg, rsg) : 0.0f) : v; };
    const int64_t xb = c * in_pitch;
    float* __restrict__ rowbase = state + chan_row_off[c] * pitch;
    uint32_t raw8[4] = {0u, 0u, 0u, 0u};  // u8: the chunk's 4 bytes
    float4a rawf[4];                      // f32: the chunk
    auto issue = [&](int64_t T) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int64_t t0 = T + 4 * (int64_t)(j * kThreads + (int)threadIdx.x);
            if (t0 < nsamps) {  // nsamps % 4 == 0: the chunk is whole
                if constexpr (IN_KIND == 1)
                    raw8[j] = *reinterpret_cast<const uint32_t*>(static_cast<const uint8_t*>(in) + xb + t0);
                else if constexpr (IN_KIND == 0)
                    rawf[j] = *reinterpret_cast<const float4a*>(static_cast<const float*>(in) + xb + t0);
            }
        }
    };
    float4a prev_last = {0.0f, 0.0f, 0.0f, 0.0f};  // thread kThreads-1: the previous tile's last chunk
    if (k0 < k1) issue(k0 * kInitTile);
    for (int64_t k = k0; k < k1; ++k) {  // block-uniform
        const int64_t T = k * kInitTile;
        float4a own[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int64_t t0 = T + 4 * (int64_t)(j * kThreads + (int)threadIdx.x);
            float4a v = {0.0f, 0.0f, 0.0f, 0.0f};
            if (t0 < nsamps) {
                if constexpr (IN_KIND == 1) {
                    const uint32_t w = raw8[j];
                    v.x = (float)(w & 0xffu);
                    v.y = (float)((w >> 8) & 0xffu);
                    v.z = (float)((w >> 16) & 0xffu);
                    v.w = (float)(w >> 24);
                } else if constexpr (IN_KIND == 0) {
                    v = rawf[j];
                } else {
                    v = {1.0f, 1.0f, 1.0f, 1.0f};  // ones: the mask transform's input
                }
                v.x = nrm(v.x);
                v.y = nrm(v.y);
                v.z = nrm(v.z);
                v.w = nrm(v.w);
            }
            own[j] = v;
        }
        if (k + 1 < k1) issue(T + kInitTile);
        __syncthreads();  // the previous tile's reads of xs are done
#pragma unroll
        for (int j = 0; j < 4; ++j) xs[1 + j * kThreads + (int)threadIdx.x] = own[j];
        if (k == k0) {
            if (threadIdx.x < 4) {
                const int64_t t = T - 4 + (int64_t)threadIdx.x;
                reinterpret_cast<float*>(xs)[threadIdx.x] = t >= 0 ? nrm(load_in<IN_KIND>(in, xb + t)) : 0.0f;
            }
        } else if (threadIdx.x == kThreads - 1) {
            xs[0] = prev_last;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int q = j * kThreads + (int)threadIdx.x;
            const int64_t t0 = T + 4 * (int64_t)q;
            if (t0 >= nsamps) continue;
            const float4a h = xs[q];  // samples t0 - 4 .. t0 - 1
            const float w[8] = {h.x, h.y, h.z, h.w, own[j].x, own[j].y, own[j].z, own[j].w};  // w[4 + e - k] = x[t0 + e - k]
            float prev[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) prev[e] = 0.0f + w[4 + e];
            if (wmask & 1u) {
                float4a o = {prev[0], prev[1], prev[2], prev[3]};
                __builtin_nontemporal_store(o, reinterpret_cast<float4a*>(rowbase + t0));
            }
#pragma unroll
            for (int kk = 1; kk < kInitConsecMax; ++kk) {
                if (kk >= K) break;  // block-uniform
                const float yk = (float)kk + 1.0f, rk = 1.0f / yk;
                const float mk = (float)(kk - 1) + 1.0f;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    float v = 0.0f;
                    if (t0 + e >= kk) v = div_rn(prev[e] * mk + (0.0f + w[4 + e - kk]), yk, rk);
                    prev[e] = v;
                }
                if ((wmask >> kk) & 1u) {
                    float4a o = {prev[0], prev[1], prev[2], prev[3]};
                    __builtin_nontemporal_store(o, reinterpret_cast<float4a*>(rowbase + (int64_t)kk * pitch + t0));
                }
            }
        }
        if (threadIdx.x == kThreads - 1) prev_last = own[3];
    }
}

// First staged pass with the mean init fused in (plans whose per-channel init
// dts are 0, 1, ..., K-1 with K <= kInitConsecMax): state 0
// never reaches HBM.  A block takes one group over tiles_per_block
// consecutive kInitGroupTile-sample tiles.  For each channel of the group the
// LDS holds its K init rows over the group's window [T - s4, T - s4 + 4 units),
// row 0 with one more 4-sample chunk of history in front; row 0 is 0 + x^ (the
// normalised sample) and row k >= 1 is (row_{k-1} * k + row0(t - k)) / (k + 1)
// (0 for t < k) -- the operations of k_init_consec in the same order (a
// division by 2 or 4 is the exact multiplication by 0.5 or 0.25).  Per tile
// every thread takes chunks of 4 samples (ChanWinD::xc numbers a channel's
// units + 1 chunks, history first): it turns a chunk's raw input, prefetched
// into registers during the previous tile's trees, into the four row-0
// values; after a barrier it builds the chunk's rows k >= 1 from row 0 in LDS
// (this chunk and the one before it); after another the trees of the group's
// rows are evaluated from LDS as in k_merge_group.  Times within a row are
// 32-bit (the host keeps nsamps below kInitGroupMaxSamps).
constexpr int kInitSlots = kInitGroupChunks / kThreads;  // chunks per thread and tile
constexpr int kInitMaxChans = 64;

struct InitChan {
    int64_t xb;       // input offset of the channel's row
    float mu, sg, rsg;
    int32_t keep;
    int32_t t_off;    // time of the channel's chunk q at T = 0: 4 q + t_off (= 4 q - s4 - 4)
    int32_t row_off;  // LDS float offset of row 0's history chunk
    int32_t rk_off;   // LDS float offset of row 1's first chunk (q = 1)
    int32_t rstride;  // 4 units: row stride of rows k >= 1
    int32_t nrows;
    int32_t xc;
};

// One whole 16-byte LDS read (p 16-byte aligned).  The empty asm keeps every
// component live, so the read is not narrowed to the components used: a
// narrowed read of each lane's 16 bytes is a strided pattern that conflicts
// in the LDS banks.
__device__ __forceinline__ float4a lds_load4(const float* p) {
    float4a v = *static_cast<const float4a*>(__builtin_assume_aligned(p, 16));
    asm volatile("" : "+v"(v));
    return v;
}

// One whole 16-byte LDS write (p 16-byte aligned; said so, or the compiler
// splits it into 4-byte pairs with the same strided pattern).
__device__ __forceinline__ void lds_store4(float* p, float4a v) {
    *static_cast<float4a*>(__builtin_assume_aligned(p, 16)) = v;
}

template <int IN_KIND>
__device__ __forceinline__ float4a unpack_chunk(uint32_t w8, const float4a& wf) {
    if constexpr (IN_KIND == 1) {
        float4a v;
        v.x = (float)(w8 & 0xffu);
        v.y = (float)((w8 >> 8) & 0xffu);
        v.z = (float)((w8 >> 16) & 0xffu);
        v.w = (float)(w8 >> 24);
        return v;
    } else {
        return wf;
    }
}

// (a - mu) / sigma correctly rounded for four samples, as div_rn computes it,
// with one branch for the rare operands outside its fast range
__device__ __forceinline__ float4a norm4(float4a v, float mu, float sg, float rsg) {
    const float a[4] = {v.x - mu, v.y - mu, v.z - mu, v.w - mu};
    float q[4];
    bool slow = !(sg >= 0x1p-60f && sg <= 0x1p60f);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const float ax = fabsf(a[e]);
        slow = slow || !(ax >= 0x1p-60f && ax <= 0x1p60f);
        const float q0 = a[e] * rsg;
        const float r = fmaf(-q0, sg, a[e]);
        q[e] = fmaf(r, rsg, q0);
    }
    if (slow) {
#pragma unroll
        for (int e = 0; e < 4; ++e) q[e] = div_rn(a[e], sg, rsg);
    }
    return float4a{q[0], q[1], q[2], q[3]};
}

// 8-bit channels whose every normalised sample lies in a range where the
// init needs no operand checks.  For v = 0..255, a(v) = fl(v - mu) and
// x(v) = fl(a(v) / sigma) are non-decreasing in v and vanish only at v = mu,
// so |a| and |x| are smallest at the integers next to mu and largest at 0 or
// 255; checking those five samples checks all 256.  If sigma and every
// nonzero |a| lie in [2^-60, 2^60], the quotient's fast form q0 = a * rsg,
// q0 + rsg * (a - q0 * sigma) is the division (norm4's fast branch); +0 for
// a = +0, never -0.  If every nonzero |x| lies in [2^-60, 2^60] too, every
// init row value is +0 or between 2^-113 and 2^64 in magnitude: sums and
// integer multiples of values on the 2^-83 grid stay on it, so the dividend
// of the division by 3 (row 2) is +0 or at least 2^-83, row 2 at least
// 2^-85 (a float on the 2^-108 grid), the dividend of the division by 5
// (row 4) +0 or at least 2^-110.  The fast quotient by the constants 3 and 5
// is then exact without checks: it scales exactly with its operand's binade
// while every step stays normal, and tests/cpp/div_check.c tries every
// mantissa of one binade against both divisors.
__device__ __forceinline__ bool u8_init_plain(float mu, float sg) {
    if (!(sg >= 0x1p-60f && sg <= 0x1p60f) || !(fabsf(mu) <= 0x1p30f)) return false;
    const int f = (int)floorf(mu);
    const int cand[5] = {0, 255, f - 1, f, f + 1};
    bool ok = true;
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        const int v = cand[i];
        if (v < 0 || v > 255) continue;
        const float a = (float)v - mu;
        if (a == 0.0f) continue;
        const float aa = fabsf(a), ax = fabsf(a / sg);
        ok = ok && aa >= 0x1p-60f && aa <= 0x1p60f && ax >= 0x1p-60f && ax <= 0x1p60f;
    }
    return ok;
}

template <int L, int IN_KIND, bool NORM, bool FAST, bool VEC>
__global__ void __launch_bounds__(kThreads) k_merge_group_init(const void* __restrict__ in, int64_t in_pitch,
                                                               ChanNorm norm, float* __restrict__ out,
                                                               int64_t out_pitch,
                                                               const GroupDescD* __restrict__ groups, int64_t ngroups,
                                                               const ChanWinD* __restrict__ chans,
                                                               const GRowDesc<L>* __restrict__ rows,
                                                               const uint8_t* __restrict__ unit_chan,
                                                               int tiles_per_block, int64_t nsamps, int64_t tile_lo,
                                                               int64_t tile_hi) {
    static_assert(IN_KIND == 0 || IN_KIND == 1, "f32 or u8 input");
    extern __shared__ __attribute__((aligned(16))) float lds[];
    __shared__ InitChan ch[kInitMaxChans];
    __shared__ int plain_all;
    const uint32_t logical = xcd_remap(blockIdx.x, gridDim.x);
    const int64_t gi = (int64_t)(logical % (uint64_t)ngroups);
    const int64_t chunk = (int64_t)(logical / (uint64_t)ngroups);
    const GroupDescD g = groups[gi];
    const int64_t k0 = tile_lo + chunk * tiles_per_block;
    const int64_t k1 = min(tile_hi, k0 + tiles_per_block);
    if (k0 >= k1) return;  // block-uniform
    if (threadIdx.x < kInitMaxChans) {  // the first wave: one lane per channel
        bool plain = true;
        if ((int)threadIdx.x < g.nleaves) {
            const ChanWinD cw = chans[g.leaf0 + threadIdx.x];
            InitChan c;
            c.xb = cw.chan * in_pitch;
            c.mu = 0.0f;
            c.sg = 1.0f;
            c.rsg = 1.0f;
            c.keep = 1;
            if (NORM) {
                c.keep = norm.keep[cw.chan] != 0;
                if (c.keep) {
                    c.mu = norm.mu[cw.chan];
                    c.sg = norm.sigma[cw.chan];
                    c.rsg = 1.0f / c.sg;
                } else {
                    c.rsg = 0.0f;  // the fast quotient of v - 0 then is +0
                }
                if (IN_KIND == 1 && c.keep) plain = u8_init_plain(c.mu, c.sg);
            }
            if (IN_KIND != 1) plain = false;
            c.t_off = -cw.s4 - 4;
            c.row_off = cw.row_off;
            c.rk_off = cw.row_off + 4 * (cw.units + 1);
            c.rstride = 4 * cw.units;
            c.nrows = cw.nrows;
            c.xc = cw.xc;
            ch[threadIdx.x] = c;
        }
        const uint64_t odd = __ballot(!plain);
        if (threadIdx.x == 0) plain_all = odd == 0;
    }
    // per slot: the chunk's channel (-1: unused) and number within it
    int32_t sc[kInitSlots], sq[kInitSlots];
#pragma unroll
    for (int j = 0; j < kInitSlots; ++j) {
        const int f = (int)threadIdx.x + j * kThreads;
        sc[j] = -1;
        sq[j] = 0;
        if (f < g.nunits) sc[j] = unit_chan[g.unit0 + f];
    }
    __syncthreads();
    // per slot: the chunk's time at T = 0, its input at T = 0, its row-0 LDS
    // offset
    int32_t qo[kInitSlots], l0[kInitSlots];
    const uint8_t* src8[kInitSlots];
    const float* srcf[kInitSlots];
#pragma unroll
    for (int j = 0; j < kInitSlots; ++j) {
        qo[j] = 0;
        l0[j] = 0;
        src8[j] = nullptr;
        srcf[j] = nullptr;
        if (sc[j] >= 0) {
            const InitChan& c = ch[sc[j]];
            sq[j] = (int)threadIdx.x + j * kThreads - c.xc;
            qo[j] = 4 * sq[j] + c.t_off;
            l0[j] = c.row_off + 4 * sq[j];
            if constexpr (IN_KIND == 1) src8[j] = static_cast<const uint8_t*>(in) + c.xb + qo[j];
            else srcf[j] = static_cast<const float*>(in) + c.xb + qo[j];
        }
    }
    const int ns = (int)nsamps;
    uint32_t raw8[kInitSlots];
    float4a rawf[kInitSlots];
#pragma unroll
    for (int j = 0; j < kInitSlots; ++j) {
        raw8[j] = 0u;
        rawf[j] = float4a{0.0f, 0.0f, 0.0f, 0.0f};
    }
    // VEC: a chunk is one aligned word (u8) or float4 (f32) load (nsamps,
    // the row pitch and the row starts multiples of 4 samples, so a chunk,
    // which starts at a multiple of 4, is whole inside the row or wholly
    // outside); otherwise four sample loads, each 0 outside [0, nsamps)
    auto issue = [&](int T) {
#pragma unroll
        for (int j = 0; j < kInitSlots; ++j) {
            if (sc[j] < 0) continue;
            const int t0 = T + qo[j];
            if constexpr (VEC) {
                if (t0 >= 0 && t0 <= ns - 4) {
                    if constexpr (IN_KIND == 1)
                        raw8[j] = *reinterpret_cast<const uint32_t*>(src8[j] + T);
                    else
                        rawf[j] = *reinterpret_cast<const float4a*>(srcf[j] + T);
                } else {
                    raw8[j] = 0u;
                    rawf[j] = float4a{0.0f, 0.0f, 0.0f, 0.0f};
                }
            } else {
                uint32_t w8 = 0u;
                float4a wf = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    if (t0 + e >= 0 && t0 + e < ns) {
                        if constexpr (IN_KIND == 1)
                            w8 |= (uint32_t)src8[j][T + e] << (8 * e);
                        else
                            wf[e] = srcf[j][T + e];
                    }
                }
                raw8[j] = w8;
                rawf[j] = wf;
            }
        }
    };
    // PLAIN: every channel of the group passed u8_init_plain (block-uniform),
    // so no operand checks, no keep select and no -0 to clear
    auto tiles = [&](auto plain_tag) {
        constexpr bool PLAIN = decltype(plain_tag)::value;
        // row 0 (0 + x^; 0 outside the row) of every chunk
        auto build0 = [&](int T) {
#pragma unroll
            for (int j = 0; j < kInitSlots; ++j) {
                if (sc[j] < 0) continue;
                const int t0 = T + qo[j];
                float4a v = unpack_chunk<IN_KIND>(raw8[j], rawf[j]);
                if constexpr (PLAIN) {
                    if (NORM) {
                        const InitChan& c = ch[sc[j]];
                        const float mu = c.mu, sg = c.sg, rsg = c.rsg;
#pragma unroll
                        for (int e = 0; e < 4; ++e) {
                            const float a = v[e] - mu;
                            const float q0 = a * rsg;
                            v[e] = fmaf(fmaf(-q0, sg, a), rsg, q0);
                        }
                    }
                } else {
                    if (NORM) {
                        const InitChan& c = ch[sc[j]];
                        v = norm4(v, c.mu, c.sg, c.rsg);
                        if (!c.keep) v = float4a{0.0f, 0.0f, 0.0f, 0.0f};
                    }
#pragma unroll
                    for (int e = 0; e < 4; ++e) v[e] = 0.0f + v[e];
                }
                float4a r0;
                if constexpr (VEC) {
                    const bool inside = t0 >= 0 && t0 <= ns - 4;
#pragma unroll
                    for (int e = 0; e < 4; ++e) r0[e] = inside ? v[e] : 0.0f;
                } else {
#pragma unroll
                    for (int e = 0; e < 4; ++e) r0[e] = (t0 + e >= 0 && t0 + e < ns) ? v[e] : 0.0f;
                }
                lds_store4(lds + l0[j], r0);
            }
        };
        // rows k >= 1 of every chunk but the history ones
        auto buildk = [&](int T) {
#pragma unroll
            for (int j = 0; j < kInitSlots; ++j) {
                if (sc[j] < 0 || sq[j] == 0) continue;
                const InitChan& c = ch[sc[j]];
                const int nrows = c.nrows;
                if (nrows < 2) continue;
                const int t0 = T + qo[j];
                const float4a h = lds_load4(lds + l0[j] - 4);
                const float4a o = lds_load4(lds + l0[j]);
                const float w[8] = {h.x, h.y, h.z, h.w, o.x, o.y, o.z, o.w};  // w[4 + e] = row 0 at t0 + e
                float* dst = lds + c.rk_off + 4 * (sq[j] - 1);
                const int rstride = c.rstride;
                // rows past the first kInitConsecMax - 1 samples need no t < k zeros
                auto rowsk = [&](auto clip_tag) {
                    constexpr bool CLIP = decltype(clip_tag)::value;
                    float prev[4] = {o.x, o.y, o.z, o.w};
#pragma unroll
                    for (int kk = 1; kk < kInitConsecMax; ++kk) {
                        if (kk >= nrows) break;
                        const float yk = (float)kk + 1.0f, rk = 1.0f / yk;
                        const float mk = (float)(kk - 1) + 1.0f;
                        const bool pow2 = kk == 1 || kk == 3;  // x / 2 and x / 4 are exact multiplications
#pragma unroll
                        for (int e = 0; e < 4; ++e) {
                            const float num = prev[e] * mk + w[4 + e - kk];
                            float val;
                            if (pow2) {
                                val = num * rk;
                            } else if (PLAIN) {
                                const float q0 = num * rk;
                                val = fmaf(fmaf(-q0, yk, num), rk, q0);
                            } else {
                                val = div_rn(num, yk, rk);
                            }
                            prev[e] = (!CLIP || t0 + e >= kk) ? val : 0.0f;
                        }
                        lds_store4(dst + (kk - 1) * rstride, float4a{prev[0], prev[1], prev[2], prev[3]});
                    }
                };
                if (t0 >= kInitConsecMax - 1) rowsk(std::false_type{});
                else rowsk(std::true_type{});
            }
        };
        issue((int)(k0 * kInitGroupTile));
        for (int64_t k = k0; k < k1; ++k) {
            const int T = (int)(k * kInitGroupTile);
            build0(T);
            __syncthreads();
            if (k + 1 < k1) issue(T + kInitGroupTile);
            buildk(T);
            __syncthreads();
            group_trees<L, false, FAST, kInitGroupTile>(lds, g, rows, out, out_pitch, T, nsamps, INT64_MAX,
                                                        SnrEpi{nullptr, 0, nullptr});
            __syncthreads();
        }
    };
    if constexpr (IN_KIND == 1) {
        if (plain_all) tiles(std::true_type{});
        else tiles(std::false_type{});
    } else {
        tiles(std::false_type{});
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
template <bool SNR, bool AUX>
static void merge_dispatch(bool vec, uint32_t nb, const float* in, int64_t in_pitch, float* out, int64_t out_pitch,
                           const RowDescD* d_rows, int64_t nrows, int64_t nsamps, int64_t limit, SnrEpi epi,
                           hipStream_t st) {
    if (vec)
        hipLaunchKernelGGL((k_merge<true, SNR, AUX>), dim3(nb), dim3(kThreads), 0, st, in, in_pitch, out, out_pitch,
                           d_rows, nrows, nsamps, limit, epi);
    else
        hipLaunchKernelGGL((k_merge<false, SNR, AUX>), dim3(nb), dim3(kThreads), 0, st, in, in_pitch, out, out_pitch,
                           d_rows, nrows, nsamps, limit, epi);
}

void launch_fdmt_merge(const float* in, int64_t in_pitch, float* out, int64_t out_pitch, const RowDescD* d_rows,
                       int64_t nrows, int64_t nsamps, int64_t out_rows_limit, hipStream_t st, const MergeEpilogue* snr,
                       bool aux) {
    if (nrows <= 0 || nsamps <= 0) return;
    const int64_t tiles = ceil_div(nsamps, kMergeTile);
    const int64_t nb = nrows * tiles;
    if (nb > 0x7fffffffLL) throw Error(SPS_ERR_INVALID_ARGUMENT, "merge grid too large");
    const bool vec = (in_pitch % 4 == 0) && (out_pitch % 4 == 0) &&
                     (reinterpret_cast<uintptr_t>(in) % 16 == 0) && (reinterpret_cast<uintptr_t>(out) % 16 == 0);
    SnrEpi epi{nullptr, 0, nullptr};
    if (snr) epi = SnrEpi{snr->nrm, snr->nrm_len, snr->stats};
    if (snr && aux) throw Error(SPS_ERR_INVALID_ARGUMENT, "no S/N epilogue on the mask transform");
    if (snr)
        merge_dispatch<true, false>(vec, (uint32_t)nb, in, in_pitch, out, out_pitch, d_rows, nrows, nsamps,
                                    out_rows_limit, epi, st);
    else if (aux)
        merge_dispatch<false, true>(vec, (uint32_t)nb, in, in_pitch, out, out_pitch, d_rows, nrows, nsamps,
                                    out_rows_limit, epi, st);
    else
        merge_dispatch<false, false>(vec, (uint32_t)nb, in, in_pitch, out, out_pitch, d_rows, nrows, nsamps,
                                     out_rows_limit, epi, st);
    SPS_HIP(hipGetLastError());
}

template <int L, bool SNR>
static void fused_dispatch2(bool vec, uint32_t nb, const float* in, int64_t in_pitch, float* out, int64_t out_pitch,
                            const void* d_trees, int64_t nrows, int64_t nsamps, int64_t limit, SnrEpi epi,
                            hipStream_t st) {
    const TreeDesc<L>* t = static_cast<const TreeDesc<L>*>(d_trees);
    if (vec)
        hipLaunchKernelGGL((k_merge_fused<L, true, SNR>), dim3(nb), dim3(kThreads), 0, st, in, in_pitch, out,
                           out_pitch, t, nrows, nsamps, limit, epi);
    else
        hipLaunchKernelGGL((k_merge_fused<L, false, SNR>), dim3(nb), dim3(kThreads), 0, st, in, in_pitch, out,
                           out_pitch, t, nrows, nsamps, limit, epi);
}

template <int L>
static void fused_dispatch(bool vec, bool snr, uint32_t nb, const float* in, int64_t in_pitch, float* out,
                           int64_t out_pitch, const void* d_trees, int64_t nrows, int64_t nsamps, int64_t limit,
                           SnrEpi epi, hipStream_t st) {
    if (snr)
        fused_dispatch2<L, true>(vec, nb, in, in_pitch, out, out_pitch, d_trees, nrows, nsamps, limit, epi, st);
    else
        fused_dispatch2<L, false>(vec, nb, in, in_pitch, out, out_pitch, d_trees, nrows, nsamps, limit, epi, st);
}

void launch_fdmt_fused(int L, const float* in, int64_t in_pitch, float* out, int64_t out_pitch, const void* d_trees,
                       int64_t nrows, int64_t nsamps, int64_t out_rows_limit, hipStream_t st,
                       const MergeEpilogue* snr) {
    if (nrows <= 0 || nsamps <= 0) return;
    const int64_t tiles = ceil_div(nsamps, kMergeTile);
    const int64_t nb = nrows * tiles;
    if (nb > 0x7fffffffLL) throw Error(SPS_ERR_INVALID_ARGUMENT, "merge grid too large");
    const bool vec = (in_pitch % 4 == 0) && (out_pitch % 4 == 0) &&
                     (reinterpret_cast<uintptr_t>(in) % 16 == 0) && (reinterpret_cast<uintptr_t>(out) % 16 == 0);
    SnrEpi epi{nullptr, 0, nullptr};
    if (snr) epi = SnrEpi{snr->nrm, snr->nrm_len, snr->stats};
    switch (L) {
        case 1:
            fused_dispatch<1>(vec, snr != nullptr, (uint32_t)nb, in, in_pitch, out, out_pitch, d_trees, nrows, nsamps,
                              out_rows_limit, epi, st);
            break;
        case 2:
            fused_dispatch<2>(vec, snr != nullptr, (uint32_t)nb, in, in_pitch, out, out_pitch, d_trees, nrows, nsamps,
                              out_rows_limit, epi, st);
            break;
        case 3:
            fused_dispatch<3>(vec, snr != nullptr, (uint32_t)nb, in, in_pitch, out, out_pitch, d_trees, nrows, nsamps,
                              out_rows_limit, epi, st);
            break;
        default:
            throw Error(SPS_ERR_INVALID_ARGUMENT, "fusion depth must be 1..3");
    }
    SPS_HIP(hipGetLastError());
}

struct GroupLaunch {
    size_t lds;
    bool onebuf;
    bool pc;  // fast tiles on the producer/consumer kernel
    const float* in;
    int64_t in_pitch;
    float* out;
    int64_t out_pitch;
    const void* groups;
    int64_t ngroups;
    const void* leaves;
    const void* rows;
    const uint8_t* unit_leaf;
    int tiles_per_block;
    int lds_stride;
    int64_t nsamps;
    int64_t limit;
    SnrEpi epi;
};

// one launch over tiles [lo, hi) of every group, tpb consecutive tiles per block
template <int L, bool SNR, bool FAST>
static void group_launch(const GroupLaunch& a, int64_t lo, int64_t hi, int tpb, hipStream_t st) {
    if (hi <= lo) return;
    const int64_t nb = a.ngroups * ceil_div(hi - lo, (int64_t)tpb);
    if (nb > 0x7fffffffLL) throw Error(SPS_ERR_INVALID_ARGUMENT, "merge grid too large");
    if (a.onebuf)
        hipLaunchKernelGGL((k_merge_group<L, SNR, FAST, true>), dim3((uint32_t)nb), dim3(kThreads), a.lds, st, a.in,
                           a.in_pitch, a.out, a.out_pitch, static_cast<const GroupDescD*>(a.groups), a.ngroups,
                           static_cast<const LeafDescD*>(a.leaves), static_cast<const GRowDesc<L>*>(a.rows),
                           a.unit_leaf, tpb, a.lds_stride, a.nsamps, a.limit, a.epi, lo, hi);
    else
        hipLaunchKernelGGL((k_merge_group<L, SNR, FAST, false>), dim3((uint32_t)nb), dim3(kThreads), a.lds, st, a.in,
                           a.in_pitch, a.out, a.out_pitch, static_cast<const GroupDescD*>(a.groups), a.ngroups,
                           static_cast<const LeafDescD*>(a.leaves), static_cast<const GRowDesc<L>*>(a.rows),
                           a.unit_leaf, tpb, a.lds_stride, a.nsamps, a.limit, a.epi, lo, hi);
}

// the producer/consumer kernel over tiles [lo, hi) (fast tiles only)
template <int L, bool SNR>
static void group_launch_pc(const GroupLaunch& a, int64_t lo, int64_t hi, int tpb, hipStream_t st) {
    if (hi <= lo) return;
    const int64_t nb = a.ngroups * ceil_div(hi - lo, (int64_t)tpb);
    if (nb > 0x7fffffffLL) throw Error(SPS_ERR_INVALID_ARGUMENT, "merge grid too large");
    const size_t bytes = 2 * sizeof(float) * (size_t)a.lds_stride;
    hipLaunchKernelGGL((k_merge_group_pc<L, SNR>), dim3((uint32_t)nb), dim3(kPCThreads), bytes, st, a.in, a.in_pitch,
                       a.out, a.out_pitch, static_cast<const GroupDescD*>(a.groups), a.ngroups,
                       static_cast<const LeafDescD*>(a.leaves), static_cast<const GRowDesc<L>*>(a.rows), a.unit_leaf,
                       tpb, a.lds_stride, a.nsamps, a.limit, a.epi, lo, hi);
}

template <int L, bool SNR>
static void pc_attributes() {
    SPS_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_merge_group_pc<L, SNR>),
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)(2 * sizeof(float) * 4 * (size_t)kGroupMaxUnits)));
}

void prepare_fdmt_group_kernels() {
    pc_attributes<1, false>();
    pc_attributes<1, true>();
    pc_attributes<2, false>();
    pc_attributes<2, true>();
    pc_attributes<3, false>();
    pc_attributes<3, true>();
    pc_attributes<4, false>();
    pc_attributes<4, true>();
}

// Tiles [fast_lo, full) take the fast-only kernel (fast_lo: the first tile at
// or past every row's largest shift and, with the S/N epilogue, the mask
// transform's extent; full: the whole tiles); the others the general one.
// Windows above kGroupTwoBufFloats run their fast tiles on the
// producer/consumer kernel when a.pc (two buffers, two 512-thread blocks per
// CU: measured 1.72 against 1.77 ms for a middle pass of the 4096 x 2^18
// search at 9216 floats, and no slower at 7168 or 8192).
template <int L, bool SNR>
static void group_dispatch(const GroupLaunch& a, int64_t fast_lo, hipStream_t st) {
    const int64_t ntiles = ceil_div(a.nsamps, kGroupTile);
    const int64_t full = a.nsamps / kGroupTile;
    if (fast_lo < full) {
        // the few first and last tiles: one per block, so that they spread
        // over the GPU instead of queueing in one block per group
        group_launch<L, SNR, false>(a, 0, fast_lo, 1, st);
        if (a.pc)
            group_launch_pc<L, SNR>(a, fast_lo, full, a.tiles_per_block, st);
        else
            group_launch<L, SNR, true>(a, fast_lo, full, a.tiles_per_block, st);
        group_launch<L, SNR, false>(a, full, ntiles, 1, st);
    } else {
        group_launch<L, SNR, false>(a, 0, ntiles, a.tiles_per_block, st);
    }
}

void launch_fdmt_group(int L, const float* in, int64_t in_pitch, float* out, int64_t out_pitch, const void* d_groups,
                       int64_t ngroups, const void* d_leaves, const void* d_rows, const uint8_t* d_unit_leaf,
                       int lds_floats_max, int tiles_per_block, int64_t nsamps, int64_t out_rows_limit,
                       int64_t max_shift, hipStream_t st, const MergeEpilogue* snr, bool pc) {
    if (ngroups <= 0 || nsamps <= 0) return;
    if (in_pitch % 4 != 0 || reinterpret_cast<uintptr_t>(in) % 16 != 0)
        throw Error(SPS_ERR_INVALID_ARGUMENT, "staged pass needs a 16-byte aligned input state");
    if (lds_floats_max <= 0 || lds_floats_max > 4 * kGroupMaxUnits)
        throw Error(SPS_ERR_INVALID_ARGUMENT, "staged pass LDS window out of range");
    if (tiles_per_block < 1) throw Error(SPS_ERR_INVALID_ARGUMENT, "tiles per block must be positive");
    if (max_shift < 0) throw Error(SPS_ERR_INVALID_ARGUMENT, "negative row shift");
    GroupLaunch a{};
    // two tile buffers while they fit 48 KiB (three blocks per CU); wider
    // windows take one buffer and a second barrier per tile (fast tiles:
    // the producer/consumer kernel)
    a.onebuf = lds_floats_max > kGroupTwoBufFloats;
    a.pc = pc && a.onebuf;
    a.lds = sizeof(float) * (a.onebuf ? 1 : 2) * (size_t)lds_floats_max;
    a.in = in;
    a.in_pitch = in_pitch;
    a.out = out;
    a.out_pitch = out_pitch;
    a.groups = d_groups;
    a.ngroups = ngroups;
    a.leaves = d_leaves;
    a.rows = d_rows;
    a.unit_leaf = d_unit_leaf;
    a.tiles_per_block = tiles_per_block;
    a.lds_stride = lds_floats_max;
    a.nsamps = nsamps;
    a.limit = out_rows_limit;
    a.epi = SnrEpi{nullptr, 0, nullptr};
    if (snr) a.epi = SnrEpi{snr->nrm, snr->nrm_len, snr->stats};
    const int64_t ext = (snr && snr->nrm_len > max_shift) ? snr->nrm_len : max_shift;
    const int64_t fast_lo = ceil_div(ext, (int64_t)kGroupTile);
    switch (L) {
        case 1:
            snr ? group_dispatch<1, true>(a, fast_lo, st) : group_dispatch<1, false>(a, fast_lo, st);
            break;
        case 2:
            snr ? group_dispatch<2, true>(a, fast_lo, st) : group_dispatch<2, false>(a, fast_lo, st);
            break;
        case 3:
            snr ? group_dispatch<3, true>(a, fast_lo, st) : group_dispatch<3, false>(a, fast_lo, st);
            break;
        case 4:
            snr ? group_dispatch<4, true>(a, fast_lo, st) : group_dispatch<4, false>(a, fast_lo, st);
            break;
        default:
            throw Error(SPS_ERR_INVALID_ARGUMENT, "staged fusion depth must be 1..4");
    }
    SPS_HIP(hipGetLastError());
}

struct GroupInitLaunch {
    size_t lds;
    const void* in;
    int64_t in_pitch;
    ChanNorm norm;
    float* out;
    int64_t out_pitch;
    const void* groups;
    int64_t ngroups;
    const void* chans;
    const void* rows;
    const uint8_t* unit_chan;
    int tiles_per_block;
    int64_t nsamps;
    bool vec;
};

template <int L, int IN_KIND, bool NORM, bool FAST>
static void group_init_launch(const GroupInitLaunch& a, int64_t lo, int64_t hi, int tpb, hipStream_t st) {
    if (hi <= lo) return;
    const int64_t nb = a.ngroups * ceil_div(hi - lo, (int64_t)tpb);
    if (nb > 0x7fffffffLL) throw Error(SPS_ERR_INVALID_ARGUMENT, "merge grid too large");
    if (a.vec)
        hipLaunchKernelGGL((k_merge_group_init<L, IN_KIND, NORM, FAST, true>), dim3((uint32_t)nb), dim3(kThreads),
                           a.lds, st, a.in, a.in_pitch, a.norm, a.out, a.out_pitch,
                           static_cast<const GroupDescD*>(a.groups), a.ngroups, static_cast<const ChanWinD*>(a.chans),
                           static_cast<const GRowDesc<L>*>(a.rows), a.unit_chan, tpb, a.nsamps, lo, hi);
    else
        hipLaunchKernelGGL((k_merge_group_init<L, IN_KIND, NORM, FAST, false>), dim3((uint32_t)nb), dim3(kThreads),
                           a.lds, st, a.in, a.in_pitch, a.norm, a.out, a.out_pitch,
                           static_cast<const GroupDescD*>(a.groups), a.ngroups, static_cast<const ChanWinD*>(a.chans),
                           static_cast<const GRowDesc<L>*>(a.rows), a.unit_chan, tpb, a.nsamps, lo, hi);
}

// the tile split of group_dispatch
template <int L, int IN_KIND, bool NORM>
static void group_init_dispatch(const GroupInitLaunch& a, int64_t fast_lo, hipStream_t st) {
    const int64_t ntiles = ceil_div(a.nsamps, kInitGroupTile);
    const int64_t full = a.nsamps / kInitGroupTile;
    if (fast_lo < full) {
        group_init_launch<L, IN_KIND, NORM, false>(a, 0, fast_lo, 1, st);
        group_init_launch<L, IN_KIND, NORM, true>(a, fast_lo, full, a.tiles_per_block, st);
        group_init_launch<L, IN_KIND, NORM, false>(a, full, ntiles, 1, st);
    } else {
        group_init_launch<L, IN_KIND, NORM, false>(a, 0, ntiles, a.tiles_per_block, st);
    }
}

template <int L>
static void group_init_kind(bool u8, bool nrm, const GroupInitLaunch& a, int64_t fast_lo, hipStream_t st) {
    if (u8)
        nrm ? group_init_dispatch<L, 1, true>(a, fast_lo, st) : group_init_dispatch<L, 1, false>(a, fast_lo, st);
    else
        nrm ? group_init_dispatch<L, 0, true>(a, fast_lo, st) : group_init_dispatch<L, 0, false>(a, fast_lo, st);
}

void launch_fdmt_group_init(int L, const void* in, int in_u8, int64_t in_pitch, const ChanNorm* norm, float* out,
                            int64_t out_pitch, const void* d_groups, int64_t ngroups, const void* d_chans,
                            const void* d_rows, const uint8_t* d_unit_chan, int lds_floats_max, int tiles_per_block,
                            int64_t nsamps, int64_t max_shift, hipStream_t st) {
    if (ngroups <= 0 || nsamps <= 0) return;
    if (in_pitch < nsamps) throw Error(SPS_ERR_INVALID_ARGUMENT, "waterfall pitch below its length");
    if (nsamps > kInitGroupMaxSamps) throw Error(SPS_ERR_INVALID_ARGUMENT, "fused init rows too long");
    const uintptr_t ia = reinterpret_cast<uintptr_t>(in);
    // whole-chunk loads need 4-sample multiples and an aligned waterfall
    const bool vec = nsamps % 4 == 0 && in_pitch % 4 == 0 && (in_u8 ? ia % 4 == 0 : ia % 16 == 0);
    if (lds_floats_max <= 0 || lds_floats_max > 16384 + 64)
        throw Error(SPS_ERR_INVALID_ARGUMENT, "staged pass LDS window out of range");
    if (tiles_per_block < 1) throw Error(SPS_ERR_INVALID_ARGUMENT, "tiles per block must be positive");
    if (max_shift < 0) throw Error(SPS_ERR_INVALID_ARGUMENT, "negative row shift");
    GroupInitLaunch a{};
    a.lds = sizeof(float) * (size_t)lds_floats_max;
    a.in = in;
    a.in_pitch = in_pitch;
    if (norm) a.norm = *norm;
    a.out = out;
    a.out_pitch = out_pitch;
    a.groups = d_groups;
    a.ngroups = ngroups;
    a.chans = d_chans;
    a.rows = d_rows;
    a.unit_chan = d_unit_chan;
    a.tiles_per_block = tiles_per_block;
    a.nsamps = nsamps;
    a.vec = vec;
    const int64_t fast_lo = ceil_div(max_shift, (int64_t)kInitGroupTile);
    const bool u8 = in_u8 != 0, nrm = norm != nullptr;
    switch (L) {
        case 1:
            group_init_kind<1>(u8, nrm, a, fast_lo, st);
            break;
        case 2:
            group_init_kind<2>(u8, nrm, a, fast_lo, st);
            break;
        case 3:
            group_init_kind<3>(u8, nrm, a, fast_lo, st);
            break;
        case 4:
            group_init_kind<4>(u8, nrm, a, fast_lo, st);
            break;
        default:
            throw Error(SPS_ERR_INVALID_ARGUMENT, "staged fusion depth must be 1..4");
    }
    SPS_HIP(hipGetLastError());
}

template <int MODE, int IN_KIND>
static void init_dispatch_norm(const void* in, int64_t in_pitch, const ChanNorm* norm, int64_t nchans,
                               int64_t nsamps, const int32_t* g, const int64_t* go, const int64_t* cro, float* state,
                               int64_t pitch, int dmax_all, const uint8_t* live, int consec, hipStream_t st) {
    ChanNorm n{};
    if (norm) n = *norm;
    if constexpr (MODE == kInitMeanD) {
        const uintptr_t ib = reinterpret_cast<uintptr_t>(in);
        const bool al_in = IN_KIND == 2 || (IN_KIND == 1 ? (ib % 4 == 0 && in_pitch % 4 == 0)
                                                         : (ib % 16 == 0 && in_pitch % 4 == 0));
        const bool al_out = (reinterpret_cast<uintptr_t>(state) % 16 == 0) && (pitch % 4 == 0);
        if (consec && al_in && al_out && nsamps % 4 == 0) {
            const int tpb = kInitTilesPerBlock;
            const dim3 grid((uint32_t)ceil_div(ceil_div(nsamps, (int64_t)kInitTile), (int64_t)tpb), (uint32_t)nchans);
            if (norm)
                hipLaunchKernelGGL((k_init_consec<IN_KIND, true>), grid, dim3(kThreads), 0, st, in, in_pitch, n,
                                   nsamps, go, cro, state, pitch, live, tpb);
            else
                hipLaunchKernelGGL((k_init_consec<IN_KIND, false>), grid, dim3(kThreads), 0, st, in, in_pitch, n,
                                   nsamps, go, cro, state, pitch, live, tpb);
            return;
        }
    }
    if (dmax_all >= 0 && dmax_all <= 16384) {
        const uintptr_t ib = reinterpret_cast<uintptr_t>(in);
        int al_in = 0;
        if (IN_KIND == 1) al_in = (ib % 16 == 0) && (in_pitch % 16 == 0) && (nsamps % 16 == 0);
        if (IN_KIND == 0) al_in = (ib % 16 == 0) && (in_pitch % 4 == 0) && (nsamps % 4 == 0);
        const int al_out = (reinterpret_cast<uintptr_t>(state) % 16 == 0) && (pitch % 4 == 0);
        const dim3 grid((uint32_t)ceil_div(nsamps, kInitTile), (uint32_t)nchans);
        const size_t lds = sizeof(float) * (size_t)(kInitTile + dmax_all);
        if (norm)
            hipLaunchKernelGGL((k_init_lds<MODE, IN_KIND, true>), grid, dim3(kThreads), lds, st, in, in_pitch, n,
                               nsamps, g, go, cro, state, pitch, dmax_all, al_in, al_out, live);
        else
            hipLaunchKernelGGL((k_init_lds<MODE, IN_KIND, false>), grid, dim3(kThreads), lds, st, in, in_pitch, n,
                               nsamps, g, go, cro, state, pitch, dmax_all, al_in, al_out, live);
        return;
    }
    dim3 grid((uint32_t)ceil_div(nsamps, kThreads), (uint32_t)nchans);
    if (norm) {
        n = *norm;
        hipLaunchKernelGGL((k_init<MODE, IN_KIND, true>), grid, dim3(kThreads), 0, st, in, in_pitch, n, nsamps, g, go,
                           cro, state, pitch, live);
    } else {
        hipLaunchKernelGGL((k_init<MODE, IN_KIND, false>), grid, dim3(kThreads), 0, st, in, in_pitch, n, nsamps, g, go,
                           cro, state, pitch, live);
    }
}

template <int MODE>
static void init_dispatch_kind(const void* in, int in_u8, int ones, int64_t in_pitch, const ChanNorm* norm,
                               int64_t nchans, int64_t nsamps, const int32_t* g, const int64_t* go,
                               const int64_t* cro, float* state, int64_t pitch, int dmax_all, const uint8_t* live,
                               int consec, hipStream_t st) {
    if (ones)
        init_dispatch_norm<MODE, 2>(in, in_pitch, norm, nchans, nsamps, g, go, cro, state, pitch, dmax_all, live,
                                    consec, st);
    else if (in_u8)
        init_dispatch_norm<MODE, 1>(in, in_pitch, norm, nchans, nsamps, g, go, cro, state, pitch, dmax_all, live,
                                    consec, st);
    else
        init_dispatch_norm<MODE, 0>(in, in_pitch, norm, nchans, nsamps, g, go, cro, state, pitch, dmax_all, live,
                                    consec, st);
}

void launch_fdmt_init(int mode, const void* in, int in_u8, int ones, int64_t in_pitch, const ChanNorm* norm,
                      int64_t nchans, int64_t nsamps, const int32_t* d_grid, const int64_t* d_grid_off,
                      const int64_t* d_chan_row_off, float* state, int64_t pitch, hipStream_t st, int dmax_all,
                      const uint8_t* d_live, int grid_consec) {
    if (nchans <= 0 || nsamps <= 0) return;
    if (nchans > 65535) throw Error(SPS_ERR_INVALID_ARGUMENT, "more than 65535 channels");
    switch (mode) {
        case kInitMeanD:
            init_dispatch_kind<kInitMeanD>(in, in_u8, ones, in_pitch, norm, nchans, nsamps, d_grid, d_grid_off,
                                           d_chan_row_off, state, pitch, dmax_all, d_live, grid_consec, st);
            break;
        case kInitRestrictD:
            init_dispatch_kind<kInitRestrictD>(in, in_u8, ones, in_pitch, norm, nchans, nsamps, d_grid, d_grid_off,
                                               d_chan_row_off, state, pitch, dmax_all, d_live, grid_consec, st);
            break;
        case kInitCFdmtD:
            init_dispatch_kind<kInitCFdmtD>(in, in_u8, ones, in_pitch, norm, nchans, nsamps, d_grid, d_grid_off,
                                            d_chan_row_off, state, pitch, dmax_all, d_live, grid_consec, st);
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
