// kernels.hpp -- host launchers for the gfx950 kernels (defined in *.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace sps {

// init formulas as seen by the kernels (same values as sps_init_mode)
constexpr int kInitMeanD = 0;
constexpr int kInitRestrictD = 1;
constexpr int kInitCFdmtD = 2;

// One merge output row: out[t] = tail[t] + (t >= offset ? head[t - offset] : 0),
// or a verbatim copy of tail when head < 0.
struct RowDescD {
    int64_t out_row;
    int64_t tail_row;
    int64_t head_row;
    int64_t offset;
};

// Per-channel normalisation applied while the FDMT init reads the waterfall:
// v = keep ? (x - mu) / sigma : 0 (or keep ? 1 : 0 for the mask transform).
struct ChanNorm {
    const float* mu;       // per channel
    const float* sigma;    // per channel
    const uint8_t* keep;   // per channel
};

// FDMT init: rows of state 0 for every channel.
//   in: nchans x nsamps (f32 or u8; ignored when ones), in_pitch elements
//   norm: nullptr, or per-channel normalisation
//   d_live: nullptr, or per state-0 row flags; rows flagged 0 are not stored
//   grid_consec: every channel's init dts are 0, 1, ..., K-1 with K <= 5
//     (enables the register fast path of the mean init)
void launch_fdmt_init(int mode, const void* in, int in_u8, int ones, int64_t in_pitch, const ChanNorm* norm,
                      int64_t nchans, int64_t nsamps, const int32_t* d_grid, const int64_t* d_grid_off,
                      const int64_t* d_chan_row_off, float* state, int64_t pitch, hipStream_t st,
                      int dmax_all = -1,  // largest init dt of any channel; < 0 selects the unstaged kernel
                      const uint8_t* d_live = nullptr, int grid_consec = 0);

// One output row of a pass that fuses L iterations: the merge tree in heap
// order (node 1 = the row; node k has tail child 2k and head child 2k+1 with
// head offset off[k], or off[k] = -1 for a copied row); leaves 2^L..2^(L+1)-1
// are rows of the pass's input state (-1 when under a missing head).
template <int L>
struct TreeDesc {
    int64_t out_row;
    int64_t leaf[1 << L];
    int32_t off[1 << L];
};
constexpr int kMaxFuse = 3;

// LDS-staged pass of L fused iterations.  Consecutive output rows are grouped
// so that the distinct input rows their trees read (the leaves) fit one LDS
// window per block: a block (group, 512-sample tile) loads every leaf row of
// the group once, over times [T - smax4, T - smax4 + nunits*4), and evaluates
// the trees of all the group's rows from LDS.
constexpr int kGroupTile = 512;
constexpr int kMaxGroupFuse = 4;
struct GroupDescD {
    int64_t row0;      // first GRowDesc of the group (live output rows, in order)
    int64_t leaf0;     // first LeafDescD of the group
    int64_t unit0;     // first entry of the group in the unit -> leaf table
    int32_t nrows;
    int32_t nleaves;
    int32_t lds_floats;
    int32_t nunits;    // float4 units of all its leaf windows (lds_floats / 4)
};
constexpr int kGroupTwoBufFloats = 6144;  // widest window that staged passes double-buffer
constexpr int kGroupMaxUnits = 2304;  // per group: 9 float4 per thread (a lone depth-4 tree needs <= 2080)
struct LeafDescD {
    int64_t row;       // input state row
    int32_t smax4;     // largest shift it is read with, rounded up to 4
    int32_t lds_off;   // float offset of its window (multiple of 4)
    int32_t nunits;    // window length in float4 units
    int32_t pad;
};
// tree of one output row with leaves as LDS bases: the leaf read at time u
// is lds[lbase + u - T] (lbase = lds_off + smax4); a leaf under a missing
// head has lbase 0 (a valid address whose value is never added)
template <int L>
struct GRowDesc {
    int64_t out_row;
    int32_t maxshift;  // largest composite shift of a leaf: for t >= maxshift every merge adds its head
    int32_t full;      // 1: no missing head (off[k] >= 0 for k = 1 .. 2^L - 1)
    int32_t lbase[1 << L];
    int32_t off[1 << L];
};

// First staged pass that computes state 0 itself (mean init, init dts 0..K-1
// per channel): for every channel whose init rows are leaves of the group,
// the K rows over [T - s4, T - s4 + 4*units) are built in LDS from the
// waterfall; row 0 (with one 4-sample history chunk in front) starts at
// row_off, row k >= 1 at row_off + 4*(units + 1) + 4*units*(k - 1).  The
// channel's units + 1 chunks are numbered from xc in the group's chunk table.
// GroupDescD.leaf0 / nleaves then index ChanWinD entries, unit0 / nunits the
// chunk -> channel table.
struct ChanWinD {
    int64_t chan;
    int64_t grid0;    // first init dt of the channel in the grid table
    int32_t nrows;    // init rows of the channel
    int32_t s4;       // largest leaf shift, rounded up to 4
    int32_t units;    // row window length / 4
    int32_t row_off;  // LDS float offset of row 0's history chunk
    int32_t xc;       // first chunk of the channel in the group's chunk numbering
    int32_t pad;
};
// The first staged pass with the init fused works on tiles of kInitGroupTile
// samples (its LDS holds every init row of the group's channels, so a shorter
// tile keeps more blocks resident: at 4096 x 2^18, 128 samples took 1.89 ms,
// 256 took 1.99 and 64 took 2.03); a group's channel windows hold at most
// kInitGroupChunks 4-sample chunks (2 per thread).
constexpr int kInitGroupTile = 128;
constexpr int kInitGroupChunks = 2 * 256;
// and times within a row are 32-bit there: rows up to 2^30 samples
constexpr int64_t kInitGroupMaxSamps = int64_t(1) << 30;

// Optional epilogue of the last merge: divide by sqrtf(n + 1e-8) with n from
// nrm (rows x nrm_len) for t < nrm_len and stats[4] (kept channels) beyond.
struct MergeEpilogue {
    const float* nrm;
    int64_t nrm_len;
    const double* stats;
};

// One merge iteration over nrows descriptors; rows with out_row >= out_rows_limit are skipped.
// aux marks launches of the mask transform (a separately named instantiation).
void launch_fdmt_merge(const float* in, int64_t in_pitch, float* out, int64_t out_pitch, const RowDescD* d_rows,
                       int64_t nrows, int64_t nsamps, int64_t out_rows_limit, hipStream_t st,
                       const MergeEpilogue* snr = nullptr, bool aux = false);

// A pass of L fused iterations (1 <= L <= kMaxFuse) over nrows TreeDesc<L>.
void launch_fdmt_fused(int L, const float* in, int64_t in_pitch, float* out, int64_t out_pitch, const void* d_trees,
                       int64_t nrows, int64_t nsamps, int64_t out_rows_limit, hipStream_t st,
                       const MergeEpilogue* snr = nullptr);

// Sets what the staged passes' producer/consumer kernels need (their LDS
// above 64 KiB); called on the current device before their first launch.
void prepare_fdmt_group_kernels();
// A staged pass (1 <= L <= kMaxGroupFuse) over ngroups GroupDescD; in must be
// 16-byte aligned with in_pitch % 4 == 0 (the internal states are).  Each
// block streams tiles_per_block consecutive tiles of one group through its
// LDS windows; d_unit_leaf maps every float4 unit of a group's windows to
// its leaf (uint8, GroupDescD.unit0 based).  max_shift: the largest
// GRowDesc::maxshift of the pass; whole tiles from there on (and past the
// S/N epilogue's nrm_len) run a kernel with the fast tree path only.
void launch_fdmt_group(int L, const float* in, int64_t in_pitch, float* out, int64_t out_pitch, const void* d_groups,
                       int64_t ngroups, const void* d_leaves, const void* d_rows, const uint8_t* d_unit_leaf,
                       int lds_floats_max, int tiles_per_block, int64_t nsamps, int64_t out_rows_limit,
                       int64_t max_shift, hipStream_t st, const MergeEpilogue* snr = nullptr, bool pc = true);

// First staged pass with the (mean) init fused in: reads the waterfall (f32 or
// u8, optional per-channel normalisation) instead of state 0; whole 4-sample
// chunk loads when nsamps, in_pitch and the waterfall's alignment allow,
// sample loads otherwise.  max_shift: as for launch_fdmt_group.
void launch_fdmt_group_init(int L, const void* in, int in_u8, int64_t in_pitch, const ChanNorm* norm, float* out,
                            int64_t out_pitch, const void* d_groups, int64_t ngroups, const void* d_chans,
                            const void* d_rows, const uint8_t* d_unit_chan, int lds_floats_max, int tiles_per_block,
                            int64_t nsamps, int64_t max_shift, hipStream_t st);

// Adjoint of one merge: g_prev[p][t] = sum over (o, s) in csr[p] of g_out[o][t+s] (t+s < nsamps).
void launch_fdmt_merge_adjoint(const float* g_out, int64_t out_pitch, float* g_prev, int64_t prev_pitch,
                               const int64_t* d_csr_off, const int64_t* d_csr_row, const int64_t* d_csr_shift,
                               int64_t prev_rows, int64_t nsamps, hipStream_t st);

// Adjoint of the (mean) init: g_x[c][i] = sum_k sum_{t=max(i,d_k)}^{min(i+d_k,N-1)} g_s[k][t] / (d_k+1).
void launch_fdmt_init_adjoint(const float* g_state, int64_t pitch, float* g_x, int64_t x_pitch, int64_t nchans,
                              int64_t nsamps, const int32_t* d_grid, const int64_t* d_grid_off,
                              const int64_t* d_chan_row_off, hipStream_t st);

// ---- statistics / normalisation ----
// row sums in FP64 (fixed lane partition) -> row mean / variance (double)
void launch_row_stats(const void* in, int in_u8, int64_t rows, int64_t cols, double* d_row_mean, double* d_row_var,
                      hipStream_t st);
// global stats + keep flags + per-channel (mu, sigma) floats; d_stats (6 doubles):
// mean, std, var_mean, var_std, n_keep, unused; d_buff (optional) gets the
// reference float layout {means[R], vars[R], mean, std, var_mean, var_std}
void launch_stats_finalize(const double* d_row_mean, const double* d_row_var, int64_t rows, int mode,
                           double* d_stats, float* d_mu, float* d_sigma, uint8_t* d_keep, float* d_buff,
                           hipStream_t st);
void launch_apply_norm(float* x, float* mask, int64_t rows, int64_t cols, const float* d_mu, const float* d_sigma,
                       const uint8_t* d_keep, hipStream_t st);
void launch_fdmt_normalization(const float* a, const float* n, int64_t len, float* out, hipStream_t st);
void launch_fdmt_normalization_rows(const float* a, int64_t rows, int64_t cols, const float* n, int64_t norm_cols,
                                    float* out, hipStream_t st);
// S/N normalisation of the DM-time image by the mask transform: for t < norm_len
// the divisor comes from d_norm (rows x norm_len), beyond it from *d_nkeep.
// The transform of a row-constant 0/1 keep mask over its first L columns,
// counted: out[r][t] = #{c : keep[c], shift[r][c] <= t} (shift: the
// accumulated delay of channel c in output row r, an int16 table; L <= 16000)
void launch_mask_counts(const int16_t* shift, const uint8_t* keep, int64_t rows, int64_t nchans, int L, float* out,
                        int64_t pitch, hipStream_t st);
void launch_snr_normalise(float* img, int64_t rows, int64_t cols, int64_t pitch, const float* d_norm,
                          int64_t norm_len, const double* d_stats, hipStream_t st);

// ---- boxcar ----
struct CandD {
    int32_t mt, mdt, mwidth;
    float msnr;
};
// Emits every (row, t, width) with window sum / sqrtf(w) > *d_thr (strict), unordered.
// pow2: widths 1,2,..,wmax (wmax a power of two), mwidth = log2 w; else widths 1..wmax, mwidth = w.
void launch_boxcar(const float* img, int64_t rows, int64_t cols, int64_t pitch, const float* d_thr, float h_thr,
                   int wmax, int pow2, CandD* out, int64_t cap, unsigned long long* d_count, hipStream_t st);
void launch_boxcar_snr(const float* img, int64_t rows, int64_t cols, int w, float* snr, hipStream_t st);
void launch_boxcar_snr_adjoint(const float* g, int64_t rows, int64_t cols, int w, float* gi, hipStream_t st);
// x[i] = floorf(x[i]) for i < n
void launch_digitize(float* x, int64_t n, hipStream_t st);
void launch_detect_max(const float* a, const float* n, int64_t rows, int64_t cols, int wnd, void* d_block_best,
                       int64_t nblocks_cap, void* d_out, hipStream_t st);
int64_t detect_max_blocks(int64_t rows, int64_t cols);
// tiled detectors of the reference's per-block kernels (mode 1: boxcar S/N of
// x / sqrtf(1e-7 + n) inside a tile; mode 2: energy ratio sqrt(sum x^2 / sum n^2));
// per-tile results in the caller's arrays of rows * ceil(cols / tile) entries
void launch_detect_tiles(int mode, const float* x, const float* n, int64_t rows, int64_t cols, int tile, int wnd,
                         float* tile_val, int32_t* tile_pos, int32_t* tile_w, void* d_out, hipStream_t st);
// zeroes the rows not kept, writes a byte mask and {row variances, mean, std of variances}
void launch_clean_rows(float* x, uint8_t* mask, int64_t rows, int64_t cols, const uint8_t* d_keep,
                       const double* d_row_var, const double* d_stats, float* d_buff, hipStream_t st);

// ---- sorting / scans / clustering ----
// stable LSD radix sort of (key, val) pairs over key bits [0, nbits); the
// result ends in (keys, vals); tmp buffers of n entries each.
void radix_sort_pairs(uint64_t* keys, uint32_t* vals, uint64_t* keys_tmp, uint32_t* vals_tmp, int64_t n, int nbits,
                      uint32_t* d_hist_scratch, int64_t hist_cap, hipStream_t st);
int64_t radix_hist_entries(int64_t n);
// inclusive scan of n uint32 values in place; scratch of scan_scratch_entries(n)
void scan_inclusive_u32(uint32_t* data, int64_t n, uint32_t* scratch, hipStream_t st);
// the same for floats: per-thread sums of 4 in order, a block tree over them,
// block prefixes added last (deterministic, not the sequential order)
void scan_inclusive_f32(float* data, int64_t n, float* scratch, hipStream_t st);
int64_t scan_scratch_entries(int64_t n);

void launch_cand_keys(const CandD* c, int64_t n, int bits_t, int bits_w, uint64_t* keys, uint32_t* vals,
                      hipStream_t st);
void launch_cand_from_keys(const uint64_t* keys, const uint32_t* vals, int64_t n, int bits_t, int bits_w,
                           CandD* out, hipStream_t st);
void launch_cluster_keys(const CandD* c, const uint32_t* perm, const uint32_t* group, int64_t n, int member,
                         uint64_t* keys, uint32_t* vals, hipStream_t st);
void launch_cluster_flags(const uint64_t* keys, int64_t n, const int32_t* d_bins, int level, uint32_t* flags,
                          hipStream_t st);
void launch_cluster_delegates(const CandD* c, const uint32_t* perm, const uint32_t* gid, const uint32_t* flags,
                              int64_t n, unsigned long long* best, uint32_t* start, hipStream_t st);
void launch_delegate_keys(const CandD* c, const unsigned long long* best, const uint32_t* start, int64_t ng,
                          int64_t n, uint64_t* keys, uint32_t* vals, hipStream_t st);
void launch_delegate_out(const unsigned long long* best, const uint32_t* start, const uint32_t* order, int64_t ng,
                         int64_t n, int32_t* delegates, int64_t* sizes, hipStream_t st);
// per caller group [begin[g], begin[g+1]) of idx: the index with the largest msnr (first on ties)
void launch_select_delegates(const CandD* c, const int32_t* idx, int64_t nidx, const int32_t* begin, int64_t ngroups,
                             int32_t* out, hipStream_t st);
void launch_iota_u32(uint32_t* v, int64_t n, hipStream_t st);
// 8-bit raw voltages -> overlapping complex FFT blocks out[pol][fft][chan][bin]
// (re/im float pairs; GUPPI input [chan][samp][npol], LOFAR [npol][samp][chan])
void launch_unpack_guppi(const int8_t* in, int64_t nsamp, int noverlap, int nbin, int nfft, int npol, int nchan,
                         float* out, hipStream_t st);
void launch_unpack_lofar(const int8_t* in, int64_t nsamp, int noverlap, int nbin, int nfft, int npol, int nchan,
                         float* out, hipStream_t st);
void launch_fill_u64(unsigned long long* v, int64_t n, unsigned long long val, hipStream_t st);

}  // namespace sps
