/*
 * sps.h -- C ABI of the MI355X single-pulse-search library (libsps.so).
 *
 * Plain C: no C++ or torch types cross this boundary.  Pointers named d_*
 * are HIP device pointers, h_* are host pointers.  Every call returns an
 * sps_status; sps_last_error() returns the message of the calling thread's
 * last failure.  Work is ordered on the stream argument (a hipStream_t passed
 * as void*, NULL = the default stream); calls that must return a host-side
 * count say so and synchronise that stream.  A handle is used by one host
 * thread at a time.
 *
 * Each entry point names the reference interface it replaces
 * (paths relative to the reference checkout).
 */
#ifndef SPS_H
#define SPS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPS_API __attribute__((visibility("default")))

typedef enum {
    SPS_OK = 0,
    SPS_ERR_INVALID_ARGUMENT = 1, /* std::invalid_argument in the reference (fdmt_base.cpp:107-117) */
    SPS_ERR_RUNTIME = 2,          /* std::runtime_error in the reference (fdmt_base.cpp:235-236,254-256) */
    SPS_ERR_HIP = 3,              /* a HIP runtime call failed */
    SPS_ERR_NO_MEMORY = 4,        /* device or host allocation failed */
    SPS_ERR_CAPACITY = 5,         /* more candidates than the output capacity (count still reported) */
    SPS_ERR_NO_DEVICE = 6         /* no HIP device is visible */
} sps_status;

typedef enum { SPS_F32 = 0, SPS_U8 = 1 } sps_dtype;

/* FDMT initialisation formulas. */
typedef enum {
    SPS_INIT_MEAN = 0,     /* running mean, FDMT_PRAVIR/fdmt_cpu.cpp:103-179 */
    SPS_INIT_RESTRICT = 1, /* prev*((dp+1)+sum)/(dc+1), FDMT_PRAVIR_RESTRICT/fdmt_cpu.cpp:65-121 */
    SPS_INIT_CFDMT = 2     /* CFdmtCpu::fnc_initC, FDMT/FdmtCpu.cpp:314-367 (legacy plans only) */
} sps_init_mode;

/* Normalisation of the waterfall before the FDMT. */
typedef enum {
    SPS_NORM_NONE = 0,
    SPS_NORM_GLOBAL = 1,     /* reference: global mean/std + outlier-variance rows zeroed
                                (CDMT/Statistical_preprocessing.cu:15-88,510-539) */
    SPS_NORM_PER_CHANNEL = 2 /* per-channel mean/std, same row cleaning */
} sps_norm_mode;

/* Candidate record, Clusterization/Candidate.cuh:4-10 (16 bytes). */
typedef struct {
    int32_t mt;     /* time sample of the window start */
    int32_t mdt;    /* DM-trial row */
    int32_t mwidth; /* log2(width) in power-of-two mode, width in all-widths mode */
    float msnr;     /* window sum / sqrtf(width) */
} sps_cand;

/* Single-maximum detector result, detection/kernel.cu:14-20. */
typedef struct {
    int32_t iwidth;
    int32_t irow;
    int32_t icol;
    float snr;
} sps_detection;

typedef struct sps_fdmt sps_fdmt;

/* ---------------------------------------------------------------------- */
/* library                                                                 */
/* ---------------------------------------------------------------------- */
SPS_API const char* sps_last_error(void);
SPS_API int sps_version(void);
/* number of visible HIP devices (0 on a host without a GPU) */
SPS_API sps_status sps_device_count(int* h_count);

/* ---------------------------------------------------------------------- */
/* plan-based FDMT: FDMT / FDMTGPU (FDMT_PRAVIR_RESTRICT/fdmt_base.hpp:37-85, */
/* fdmt_gpu.cuh:68-89).  Plan building is host-only; device state is       */
/* allocated on first use.                                                 */
/* ---------------------------------------------------------------------- */

/* FDMT::FDMT(f_min, f_max, nchans, nsamps, tsamp, dt_max, dt_step, dt_min)
 * (fdmt_base.cpp:37-53) + configure_fdmt_plan (fdmt_base.cpp:119-279). */
SPS_API sps_status sps_fdmt_create(float f_min, float f_max, int64_t nchans, int64_t nsamps, float tsamp,
                                   int64_t dt_max, int64_t dt_step, int64_t dt_min, int init_mode,
                                   sps_fdmt** out);

/* CFdmtGpu::CFdmtGpu(Fmin, Fmax, nchan, cols, imaxDt) (FDMT/FdmtGpu.cu:230-270,
 * config from FDMT/FdmtCpu.cpp:143-194). */
SPS_API sps_status sps_cfdmt_create(float f_min, float f_max, int nchan, int cols, int imaxdt, sps_fdmt** out);

SPS_API sps_status sps_fdmt_destroy(sps_fdmt* h);

/* FDMT::get_niters (fdmt_base.cpp:60); CFdmtB m_iNumIter */
SPS_API sps_status sps_fdmt_get_niters(const sps_fdmt* h, int64_t* h_niters);
/* plan.state_shape[i] = {nchans, ndt_min, ndt_max, sum ndt, nsamps} (fdmt_base.hpp:10-11) */
SPS_API sps_status sps_fdmt_get_state_shape(const sps_fdmt* h, int64_t iter, int64_t* h_shape5);
/* rows of the caller's output (dmt) and samples per row */
SPS_API sps_status sps_fdmt_get_output_shape(const sps_fdmt* h, int64_t* h_rows, int64_t* h_cols);
/* FDMT::get_dt_grid_final (fdmt_base.cpp:62-64); h_n receives the length */
SPS_API sps_status sps_fdmt_get_dt_grid_final(const sps_fdmt* h, int64_t* h_out, int64_t cap, int64_t* h_n);
/* FDMT::get_dm_grid_final (fdmt_base.cpp:65-75) */
SPS_API sps_status sps_fdmt_get_dm_grid_final(const sps_fdmt* h, float* h_out, int64_t cap, int64_t* h_n);
/* FDMT::get_df / get_correction (fdmt_base.cpp:56-57) */
SPS_API sps_status sps_fdmt_get_df(const sps_fdmt* h, float* h_df, float* h_correction);
/* plan tables of iteration iter (FDMTPlan, fdmt_base.hpp:20-35):
 * h_grid_off (nsub+1) and h_grid (sum ndt) for dt_grid[iter];
 * h_map7 rows {out_sub,out_dt,head_sub,head_dt,tail_sub,tail_dt,offset},
 * h_copy4 rows {out_sub,out_dt,tail_sub,tail_dt}; any pointer may be NULL to
 * query only the counts. */
SPS_API sps_status sps_fdmt_get_plan_iter(const sps_fdmt* h, int64_t iter, int64_t* h_nsub, int64_t* h_grid_off,
                                          int64_t* h_grid, int64_t* h_nmap, int64_t* h_map7, int64_t* h_ncopy,
                                          int64_t* h_copy4, float* h_df_top, float* h_df_bot);
/* lowered engine rows of iteration iter (>= 1): h_rows4 {out,tail,head(-1 copy),offset} */
SPS_API sps_status sps_fdmt_get_rows(const sps_fdmt* h, int64_t iter, int64_t* h_n, int64_t* h_rows4);

/* FDMT::execute(waterfall, waterfall_size, dmt, dmt_size) (fdmt_base.hpp:55-56,
 * FDMTGPU::execute fdmt_gpu.cu:235-347) and CFdmtGpu::process_image(in, out,
 * b_ones) (FdmtGpu.cu:273-348).  d_waterfall: nchans x nsamps (dtype f32 or
 * u8 promoted to f32); d_dmt: output rows x nsamps.  Sizes are checked like
 * check_inputs (fdmt_base.cpp:107-117).  ones != 0 runs the transform of an
 * all-ones input (d_waterfall may be NULL). */
SPS_API sps_status sps_fdmt_execute(sps_fdmt* h, const void* d_waterfall, int dtype, size_t waterfall_size,
                                    float* d_dmt, size_t dmt_size, int ones, void* stream);

/* Merge schedule (no reference counterpart; a tuning knob): the merge
 * iterations run as passes that each fuse 1..3 consecutive iterations into one
 * launch (same fp32 operations and results; intermediate states stay out of
 * HBM).  h_depths[k] is the depth of pass k and must sum to niters. */
SPS_API sps_status sps_fdmt_set_schedule(sps_fdmt* h, const int32_t* h_depths, int64_t n);
SPS_API sps_status sps_fdmt_get_schedule(const sps_fdmt* h, int32_t* h_depths, int64_t cap, int64_t* h_n);
/* Staged passes (no reference counterpart; a tuning knob): each pass groups
 * consecutive output rows so that the input rows their merge trees read fit
 * one LDS window per block, reads each of those rows once per time tile and
 * evaluates the trees from LDS (same fp32 operations and results).  Staged
 * passes fuse 1..4 iterations; turning staging off resets a schedule with
 * depth-4 passes to the default. */
SPS_API sps_status sps_fdmt_set_staged(sps_fdmt* h, int staged);
/* Shape of the staged passes: LDS window per tile buffer (2048..9216 floats,
 * default 8192; a block holds two), output rows per group (1..64, default
 * 16) and consecutive 512-sample tiles streamed per block (1..512, default 8). */
SPS_API sps_status sps_fdmt_set_group_shape(sps_fdmt* h, int lds_floats, int max_rows, int tiles_per_block);

/* FDMTCPU::execute (FDMT_PRAVIR_RESTRICT/fdmt_cpu.hpp:5-20, fdmt_cpu.cpp:28-54)
 * and CFdmtCpu::process_image (FDMT/FdmtCpu.h:6-34, FdmtCpu.cpp:197-249):
 * the same transform as sps_fdmt_execute on HOST buffers (f32 waterfall,
 * nchans x nsamps; dmt, output rows x nsamps).  The work runs on the GPU: the
 * waterfall is copied in, transformed and the result copied back before the
 * call returns (default stream, synchronous).  Sizes are checked like
 * check_inputs.  ones != 0: transform of an all-ones input (h_waterfall may
 * be NULL), the b_ones mode of process_image. */
SPS_API sps_status sps_fdmt_execute_host(sps_fdmt* h, const float* h_waterfall, size_t waterfall_size, float* h_dmt,
                                         size_t dmt_size, int ones);
/* FDMTCPU::initialise (fdmt_cpu.hpp:13) on host buffers: the iteration-0
 * state (rows0 x nsamps) of a host waterfall, computed on the GPU. */
SPS_API sps_status sps_fdmt_initialise_host(sps_fdmt* h, const float* h_waterfall, float* h_state,
                                            size_t state_size);
/* Legacy plan configuration of a sps_cfdmt_create handle (CFdmtCpu::
 * create_config, FDMT/FdmtCpu.cpp:143-194; the members m_parrQuantMtrx_h,
 * m_pparrRowsCumSum_h, m_pparrFreq_h of FDMT/FdmtCpu.h:21-24): for iteration
 * iter (0..niters), h_quant receives the sub-matrix count q, and
 * h_rows_cumsum / h_freq (q + 1 entries each, may be NULL) the cumulative
 * row offsets and the sub-band edge frequencies. */
SPS_API sps_status sps_cfdmt_get_config(const sps_fdmt* h, int64_t iter, int32_t* h_quant, int32_t* h_rows_cumsum,
                                        float* h_freq);

/* FDMT::initialise (fdmt_base.hpp:57): the iteration-0 state, rows0 x nsamps */
SPS_API sps_status sps_fdmt_initialise(sps_fdmt* h, const void* d_waterfall, int dtype, float* d_state,
                                       size_t state_size, void* stream);

/* Adjoint (transpose) of execute: d_grad_waterfall (nchans x nsamps) =
 * FDMT^T d_grad_dmt (rows x nsamps).  No reference counterpart (used by the
 * differentiable search path); linear, so no forward state is kept. */
SPS_API sps_status sps_fdmt_adjoint(sps_fdmt* h, const float* d_grad_dmt, size_t dmt_size, float* d_grad_waterfall,
                                    size_t waterfall_size, void* stream);

/* ---------------------------------------------------------------------- */
/* normalisation                                                           */
/* ---------------------------------------------------------------------- */

/* statistical_preprocessing::calc_fdmt_inputs(d_in, NRows, NCols, d_buff,
 * d_normInput) (CDMT/Statistical_preprocessing.cuh:9-10, .cu:15-88).
 * d_in (f32, in place) is cleaned and normalised; d_norm_input (may be NULL)
 * receives the 0/1 mask; d_buff (>= 2*nrows+4 floats) receives
 * {row means, row variances, mean, std, var mean, var std} like the
 * reference's layout.  mode: sps_norm_mode (GLOBAL = reference). */
SPS_API sps_status sps_calc_fdmt_inputs(float* d_in, int64_t nrows, int64_t ncols, float* d_buff,
                                        float* d_norm_input, int mode, void* stream);

/* fdmt_normalization(d_arr, d_norm, len, d_out) (CDMT/Chunk_gpu.cu:841-852):
 * d_out[i] = d_arr[i] / sqrtf(d_norm[i] + 1e-8); d_out may alias d_arr. */
SPS_API sps_status sps_fdmt_normalization(const float* d_arr, const float* d_norm, int64_t len, float* d_out,
                                          void* stream);
/* The same division over a rows x cols image whose norm image is causal and
 * constant along each row from column norm_cols - 1 on (the transform of a
 * row-constant keep mask is, once t passes the largest shift), so only its
 * first norm_cols columns are given: d_out[r][t] = d_arr[r][t] /
 * sqrtf(d_norm[r][min(t, norm_cols - 1)] + 1e-8), d_norm rows x norm_cols
 * (no reference counterpart: it lets CChunk_gpu::process transform a
 * narrow slice of the mask; d_out may alias d_arr). */
SPS_API sps_status sps_fdmt_normalization_rows(const float* d_arr, int64_t rows, int64_t cols, const float* d_norm,
                                               int64_t norm_cols, float* d_out, void* stream);

/* ---------------------------------------------------------------------- */
/* boxcar candidate gathering                                              */
/* ---------------------------------------------------------------------- */

/* clusterization::gather_candidates_in_dynamicalArray (Clusterization.cuh:22-23,
 * .cu:56-89,419-503): widths 1,2,..,2^q, mwidth = log2(width).
 * gather_candidates_in_heap_with_plan (Clusterization.cuh:56-63, .cu:169-202,
 * 591-671): all_widths != 0 gives widths 1..wnd with mwidth = width.
 * d_thr: device pointer to the threshold (the reference passes a reference
 * to device memory).  Candidates are written in (mdt, mt, mwidth) order;
 * h_count receives the total (synchronises the stream); more than cap
 * returns SPS_ERR_CAPACITY with the first cap candidates in order. */
SPS_API sps_status sps_boxcar_candidates(const float* d_img, int64_t rows, int64_t cols, const float* d_thr,
                                         int wnd, int all_widths, sps_cand* d_out, int64_t cap, int64_t* h_count,
                                         void* stream);

/* clusterization::complete_clusterization (Clusterization.cuh:25-32, .cu:92-166)
 * and select_delegates (select_delegates.cu:62-125): three-key gap clustering
 * of d_cands with d_bins (device, 3 ints for mt, mdt, mwidth); d_delegates
 * (n entries) receives delegate candidate indices sorted by mt and
 * d_group_size their group sizes; h_ngroups the group count (synchronises). */
SPS_API sps_status sps_cluster(const sps_cand* d_cands, int64_t n, const int32_t* d_bins, int32_t* d_delegates,
                               int64_t* d_group_size, int64_t* h_ngroups, void* stream);

/* doDetectSignal_v1 single-maximum detector (detection/kernel.cu:561-582):
 * max over rows, windows 1..wnd fully inside, of the boxcar S/N of
 * d_arr / sqrtf(1e-7 + d_norm) (d_norm may be NULL: no division).  Ties go to
 * the smaller width, then row, then column. */
SPS_API sps_status sps_detect_max(const float* d_arr, const float* d_norm, int64_t rows, int64_t cols, int wnd,
                                  sps_detection* d_out, void* stream);

/* select_delegates::select_delegates / select_delegates_kernel
 * (select_delegates.cuh:4-14, select_delegates.cu:16-125) over groups the
 * caller made: group g is positions [d_group_begin[g], d_group_begin[g+1])
 * of d_indices (the last ends at nidx); d_delegates[g] receives the
 * candidate index there with the largest msnr, the first such position on
 * ties (thrust::max_element's rule; a NaN at a group's first position is
 * kept as max_element keeps it); -1 for an empty group. */
SPS_API sps_status sps_select_delegates(const sps_cand* d_cands, const int32_t* d_indices, int64_t nidx,
                                        const int32_t* d_group_begin, int64_t ngroups, int32_t* d_delegates,
                                        void* stream);

/* select_delegates::write_log (select_delegates.cuh:15-25,
 * select_delegates.cu:127-177): writes filename with the columns
 * N, t bin, t sec (t_chunk_begin + mt * delt_bin), dedisp, width, SNR, num
 * (each std::setw(10)) for the delegates h_delegates[0..ndel) (indices into
 * the ncand device candidates) and their group sizes h_group_sizes. */
SPS_API sps_status sps_write_log(const sps_cand* d_cands, int64_t ncand, const int32_t* h_delegates,
                                 const int64_t* h_group_sizes, int64_t ndel, const char* filename, float delt_bin,
                                 float t_chunk_begin);

/* Inclusive prefix sums over n values into d_out (may equal d_in):
 * fnc_cumsum_customized_gpu / calc_cumsum_custom_gpu and inclusiveScan_CUB
 * (Clusterization/cumsum_customized.cu:147-179,259-305).  int32 sums wrap
 * like two's-complement addition (bit-exact with any order); float sums are
 * formed in a fixed blocked order (4 values per thread, a tree over the
 * block, block prefixes added last), deterministic but not the reference's
 * CUB order: equal to the sequential sums within the error of a sum of n
 * terms.  The scratch is allocated and released inside; synchronises. */
SPS_API sps_status sps_scan_inclusive_i32(const int32_t* d_in, int32_t* d_out, int64_t n, void* stream);
SPS_API sps_status sps_scan_inclusive_f32(const float* d_in, float* d_out, int64_t n, void* stream);

/* 8-bit raw-voltage unpack into overlapping complex FFT blocks, the input
 * side of the coherent front-end: d_out[pol][fft][chan][bin] (pol < npol/2,
 * re/im float pairs, nfft * nchan * nbin * npol complex-halves in all) takes
 * sample isamp = bin + (nbin - 2*noverlap)*fft - noverlap of the chunk, 0 + 0i
 * outside [0, nsamp); signed bytes promote exactly to float.  npol is 2 or 4
 * (re/im per polarisation), nbin > 2*noverlap.
 *   sps_unpack_guppi: d_in[chan][nsamp][npol] -- CSession_guppi_gpu::
 *     unpack_chunk (CDMT/Session_guppi_gpu.cu:53-60) and its kernel
 *     unpack_chunk_guppi_gpu (:116-140; unpack_guppi/kernel.cu:161-187);
 *   sps_unpack_lofar: d_in[npol][nsamp][nchan] (component 2p is pol p's
 *     real part, 2p + 1 its imaginary part) -- CSession_lofar_gpu::
 *     unpack_chunk (CDMT/Session_lofar_gpu.cu:54-67) and unpackInput_lofar
 *     (:144-175; unpack_guppi/kernel.cu:81-122).
 * Asynchronous on the stream. */
SPS_API sps_status sps_unpack_guppi(const int8_t* d_in, int64_t nsamp, int noverlap, int nbin, int nfft, int npol,
                                    int nchan, float* d_out, void* stream);
SPS_API sps_status sps_unpack_lofar(const int8_t* d_in, int64_t nsamp, int noverlap, int nbin, int nfft, int npol,
                                    int nchan, float* d_out, void* stream);

/* Coherent front-end between the unpack and the search: CChunk_gpu::process
 * with CChunk_fly_gpu's on-the-fly chirp (CDMT/Chunk_gpu.cu:197-209 plans,
 * :254-322 steps 2-8; Chunk_fly_gpu.cu:94-164 chirp).  A handle holds the
 * two rocFFT plans of CChunk_gpu::create_fft_plans (forward: nbin points,
 * nfft * nchan * npol/2 blocks; inverse: mbin = nbin / len_sft points, len_sft
 * times as many), the sub-band taper per bin and,
 * on the unfused path only, one scratch chunk.  Spectra are complex float
 * [pol][fft][chan][bin] (the unpack's layout, npol/2 polarisations); the
 * waterfall is float [nchan * len_sft][msamp], msamp = nfft * (mbin - 2 *
 * (noverlap / len_sft)), rows in the reference's order (both frequency axes
 * reversed), each row shifted by its incoherent dispersion delay.  All calls
 * are asynchronous on the stream. */
typedef struct sps_coherent sps_coherent;
SPS_API sps_status sps_coherent_create(float f_min, float f_max, int npol, int nchan, int len_sft, int nbin,
                                       int nfft, int noverlap, float tsamp, sps_coherent** out);
SPS_API sps_status sps_coherent_destroy(sps_coherent* h);
/* waterfall rows and columns, and the complex values of one chunk */
SPS_API sps_status sps_coherent_get_shape(const sps_coherent* h, int64_t* h_rows, int64_t* h_cols,
                                          int64_t* h_spectra);
/* *h_fused = 1 when sps_coherent_waterfall takes the fused pass (no
 * reference counterpart: a property of this implementation) */
SPS_API sps_status sps_coherent_is_fused(const sps_coherent* h, int* h_fused);
/* step 2: cufftExecC2C(m_fftPlanForward) in place (Chunk_gpu.cu:267) */
SPS_API sps_status sps_coherent_forward(sps_coherent* h, float* d_blocks, void* stream);
/* steps 4-8 for one coherent DM into d_waterfall.  When the sub-band length
 * mbin is a power of two in 128..4096 and len_sft is even, one fused pass per
 * DM does the chirp, a length-mbin inverse FFT in LDS and the intensity
 * waterfall, reading the spectra once (results within float FFT rounding of
 * the step-by-step path); otherwise chirp into the handle's scratch chunk,
 * inverse FFT there (rocFFT) and the intensity pass. */
SPS_API sps_status sps_coherent_waterfall(sps_coherent* h, const float* d_spectra, double dm, float* d_waterfall,
                                          void* stream);
/* the steps one at a time: the chirp (kernel_el_wise_mult_onthe_fly,
 * Chunk_fly_gpu.cu:119-164) from d_spectra into d_out; the inverse FFT in
 * place (Chunk_gpu.cu:288); the intensity waterfall from inverse-FFT output
 * (roll_rows_normalize_sum_kernel, transpose_unpadd_intensity and
 * dedisperse, Chunk_gpu.cu:476-515,613-632, as one pass) */
SPS_API sps_status sps_coherent_chirp(sps_coherent* h, const float* d_spectra, double dm, float* d_out,
                                      void* stream);
SPS_API sps_status sps_coherent_inverse(sps_coherent* h, float* d_buf, void* stream);
SPS_API sps_status sps_coherent_detect(sps_coherent* h, const float* d_buf, double dm, float* d_waterfall,
                                       void* stream);

/* statistical_preprocessing::clean_fdmtInput_and_create_normInput
 * (CDMT/Statistical_preprocessing.cuh:6-7, .cu:102-120,142-268): rows whose
 * variance lies more than 4 standard deviations from the mean row variance
 * are zeroed in place (values are not normalised); d_mask (may be NULL)
 * receives one byte per sample, 1 where the row is kept, 0 where it was
 * cleaned; d_buff (may be NULL, >= nrows + 2 floats) receives the row
 * variances, their mean and their standard deviation.  The reference never
 * writes its mask argument (and reads its device-resident mean and std on
 * the host); this entry point fills the mask the function's name promises.
 * Synchronises the stream. */
SPS_API sps_status sps_clean_fdmt_input(float* d_arr, int64_t nrows, int64_t ncols, uint8_t* d_mask, float* d_buff,
                                        void* stream);

/* doDetectSignal_v1 / doDetectSignal_v2 (detection/kernel.cu:561-605) with
 * the window arithmetic of their per-block kernels (:42-226): rows are cut
 * into tiles of `tile` columns (the reference's block width), inside which
 * fp32 prefix sums run sequentially from the tile's first column.
 *   version 1: value = (Y[t+w] - Y[t]) / sqrtf(w), y = x / sqrtf(1e-7 + n)
 *              (n may be NULL: y = x); windows lie inside their tile.
 *   version 2: value = sqrtf((X2[t+w]-X2[t]) / (N2[t+w]-N2[t])), X2, N2
 *              prefix sums of x^2 and n^2 over the tile and the wnd samples
 *              after it, N2[0] = 1e-6 (n may be NULL: n = 1).
 * Widths 1..wnd; windows end inside the row.  The per-tile best goes to the
 * caller's arrays d_tile_val / d_tile_pos (row*cols + column) / d_tile_w
 * (the reference's d_pAuxArray, d_pAuxNumArray, d_pWidthArray; at least
 * rows * ceil(cols/tile) entries, ntiles_cap), and the global best to d_out.
 * Ties: larger value, then smaller width, row, column (the reference's tie
 * order follows its tree reductions and is not reproduced).  Allocates
 * nothing. */
SPS_API sps_status sps_detect_signal(int version, const float* d_arr, const float* d_norm, int64_t rows, int64_t cols,
                                     int wnd, int tile, float* d_tile_val, int32_t* d_tile_pos, int32_t* d_tile_w,
                                     int64_t ntiles_cap, sps_detection* d_out, void* stream);

/* No counterpart in the reference: sps_detect_max (the same arguments and
 * the same result, bit for bit) for callers that cannot allocate device
 * memory -- code under stream capture, tight serving loops, harnesses that
 * forbid hipMalloc.  Its per-block partial results live in a static device
 * array of 4096 x 16 bytes per GPU instead of a buffer allocated per call,
 * so it calls no allocator; any image size is accepted (rows and columns
 * up to 2^31 - 1).  Calls on one GPU share that array: calls that overlap
 * in time (different streams or threads) race; calls ordered on one stream
 * do not. */
SPS_API sps_status sps_detect_max_noalloc(const float* d_arr, const float* d_norm, int64_t rows, int64_t cols, int wnd,
                                          sps_detection* d_out, void* stream);

/* clusterization::digitize_kernel (Clusterization.cuh:34-35,
 * Clusterization.cu:206-214): d_x[i] = floorf(d_x[i]) in place, i < n (the
 * reference test harness uses it to make integer-valued images). */
SPS_API sps_status sps_digitize(float* d_x, int64_t n, void* stream);

/* Boxcar S/N map and its adjoint for the differentiable path (no reference
 * counterpart): d_snr[r][t] = sum_{j<w} d_img[r][t+j] / sqrtf(w) for t+w <= cols
 * (0 elsewhere); adjoint d_grad_img[r][t'] = sum_{t<=t'<t+w} d_grad_snr[r][t]/sqrtf(w). */
SPS_API sps_status sps_boxcar_snr(const float* d_img, int64_t rows, int64_t cols, int w, float* d_snr, void* stream);
SPS_API sps_status sps_boxcar_snr_adjoint(const float* d_grad_snr, int64_t rows, int64_t cols, int w,
                                          float* d_grad_img, void* stream);

/* ---------------------------------------------------------------------- */
/* fused single-beam search: the CChunk_gpu::process hot section           */
/* (CDMT/Chunk_gpu.cu:325-425): normalise -> FDMT -> S/N normalisation ->  */
/* boxcar gather, without host synchronisation.                            */
/* ---------------------------------------------------------------------- */
typedef struct sps_search sps_search;

typedef struct {
    float f_min, f_max, tsamp;
    int64_t nchans, nsamps, dt_max;
    int in_dtype;      /* sps_dtype */
    int norm_mode;     /* sps_norm_mode */
    int snr_norm;      /* divide by sqrt(FDMT(mask) + 1e-8) like fdmt_normalization */
    int box_q;         /* widths 1..2^box_q */
    float threshold;   /* strict > */
    int64_t cand_capacity;
} sps_search_params;

SPS_API sps_status sps_search_create(const sps_search_params* p, sps_search** out);
/* The same search over the legacy CFdmt plan, as CChunk_gpu::process runs
 * it for each coherent DM (m_Fdmt = CFdmtGpu(Fmin, Fmax, nchan * len_sft,
 * msamp, len_sft), Chunk_gpu.cu:175-181; statistics, the transforms of mask
 * and waterfall and fdmt_normalization, :325-356): p->dt_max and p->tsamp
 * are unused, the image has imaxdt rows. */
SPS_API sps_status sps_search_create_cfdmt(const sps_search_params* p, int imaxdt, sps_search** out);
SPS_API sps_status sps_search_destroy(sps_search* s);
/* Runs one beam.  d_in: nchans x nsamps; d_dmt (may be NULL): the S/N
 * normalised DM-time image (ndt x nsamps); d_cands (cand_capacity) and
 * d_count (one uint64, total before capacity clipping) are written in
 * stream order; candidates are in no particular order until
 * sps_search_sort_candidates. */
SPS_API sps_status sps_search_run(sps_search* s, const void* d_in, float* d_dmt, sps_cand* d_cands,
                                  uint64_t* d_count, void* stream);
/* Sorts the first min(count, capacity) candidates by (mdt, mt, mwidth) and
 * returns the count (synchronises).  A count above cand_capacity returns
 * SPS_ERR_CAPACITY after sorting the kept prefix and setting h_count: which
 * candidates were kept depends on the order the device counter was taken,
 * so the caller re-gathers into a buffer of h_count entries with
 * sps_search_gather (the reference's gather counts first and then
 * allocates exactly, Clusterization.cu:56-88, so it never drops one). */
SPS_API sps_status sps_search_sort_candidates(sps_search* s, sps_cand* d_cands, const uint64_t* d_count,
                                              int64_t* h_count, void* stream);
/* Re-runs only the boxcar gather of the last sps_search_run over its S/N
 * image (d_dmt: the image that run wrote, or NULL when the run kept it
 * internally) into d_cands (cap entries), sorted by (mdt, mt, mwidth);
 * d_count and h_count receive the total (synchronises); SPS_ERR_CAPACITY
 * when it exceeds cap. */
SPS_API sps_status sps_search_gather(sps_search* s, const float* d_dmt, sps_cand* d_cands, int64_t cap,
                                     uint64_t* d_count, int64_t* h_count, void* stream);
/* The clustering leg of CChunk_gpu::process (CDMT/Chunk_gpu.cu:398-425:
 * clusterization_main, Clusterization.cu:17-54) on the S/N image of the
 * last sps_search_run: the all-widths gather (widths 1..wnd, mwidth = width,
 * the search's threshold, strict >) over d_dmt (the image that run wrote, or
 * NULL when it kept the image internally), the three-key gap clustering
 * with d_bins (device, 3 ints for mt, mdt, mwidth) and the delegates (the
 * largest S/N of each group, sorted by mt).  h_ncand receives the number of
 * gathered candidates and h_ngroups the number of groups (synchronises). */
SPS_API sps_status sps_search_cluster(sps_search* s, const float* d_dmt, int wnd, const int32_t* d_bins,
                                      int64_t* h_ncand, int64_t* h_ngroups, void* stream);
/* The delegates of the last sps_search_cluster: up to cap candidate records
 * (sorted by mt) and their group sizes (either pointer may be NULL); h_n
 * receives the group count. */
SPS_API sps_status sps_search_get_delegates(const sps_search* s, sps_cand* h_delegates, int64_t* h_group_sizes,
                                            int64_t cap, int64_t* h_n);
/* select_delegates::write_log (select_delegates.cu:127-177) of the last
 * sps_search_cluster's delegates, as sps_write_log writes it. */
SPS_API sps_status sps_search_write_log(const sps_search* s, const char* filename, float delt_bin,
                                        float t_chunk_begin);
SPS_API sps_status sps_search_get_ndt(const sps_search* s, int64_t* h_ndt);
/* staged passes of the search's data transform (see sps_fdmt_set_staged) */
SPS_API sps_status sps_search_set_staged(sps_search* s, int staged);
SPS_API sps_status sps_search_set_group_shape(sps_search* s, int lds_floats, int max_rows, int tiles_per_block);
/* merge schedule of the search's data transform (see sps_fdmt_set_schedule) */
SPS_API sps_status sps_search_set_schedule(sps_search* s, const int32_t* h_depths, int64_t n);
SPS_API sps_status sps_search_get_schedule(const sps_search* s, int32_t* h_depths, int64_t cap, int64_t* h_n);
/* Stage timing (no reference counterpart; a profiling aid).  When enabled,
 * sps_search_run records HIP events on its stream between stages; after a run,
 * sps_search_get_timing synchronises and writes the stage durations in ms:
 * [normalise stats, mask transform, FDMT init, merge pass 1 .. merge pass P
 *  (the last one carries the fused S/N division), separate S/N
 *  normalisation (zero when fused), boxcar] (h_n receives P + 5, P = the
 *  number of passes of the schedule). */
SPS_API sps_status sps_search_enable_timing(sps_search* s, int enable);
SPS_API sps_status sps_search_get_timing(sps_search* s, double* h_ms, int64_t cap, int64_t* h_n);
/* Compulsory HBM bytes of each stage in the order of sps_search_get_timing:
 * the statistics' input read (at the input's element size); the mask
 * transform's output image written and read once; the init's input read and
 * state-0 write; each pass's read of its input state and write of its output
 * state; the boxcar's read of the image.  States count only their live rows
 * (rows some output row depends on; the library computes only those). */
SPS_API sps_status sps_search_get_stage_bytes(const sps_search* s, double* h_bytes, int64_t cap, int64_t* h_n);

/* ---------------------------------------------------------------------- */
/* batched multi-beam search (no reference counterpart: the reference runs  */
/* one chunk per CChunk_gpu::process call, CDMT/Chunk_gpu.cu:325-425).      */
/* Many independent beams of any channel count and length run the search    */
/* above in one launch sequence: beams are spread over nworkers streams     */
/* (largest first onto the least loaded), the beams of one stream share one  */
/* set of scratch states, and sps_batch_launch captures the whole fork/join */
/* sequence into a HIP graph once and replays it on later calls.            */
/* ---------------------------------------------------------------------- */
typedef struct sps_batch sps_batch;

typedef struct {
    float f_min, f_max, tsamp;
    int64_t nchans, nsamps, dt_max;
    const void* d_in;       /* nchans x nsamps input of the batch's dtype */
    float* d_dmt;           /* ndt x nsamps S/N image, or NULL (kept in scratch) */
    sps_cand* d_cands;      /* cand_capacity candidates */
    int64_t cand_capacity;
    uint64_t* d_count;      /* candidates found, before capacity clipping */
} sps_batch_beam;

typedef struct {
    int in_dtype;      /* sps_dtype, every beam */
    int norm_mode;     /* sps_norm_mode */
    int snr_norm;
    int box_q;
    float threshold;
    int nworkers;      /* concurrent streams, 1..8 */
} sps_batch_params;

/* The buffers named in beams are bound to the batch (and to its graph). */
SPS_API sps_status sps_batch_create(const sps_batch_params* p, const sps_batch_beam* beams, int64_t nbeams,
                                    sps_batch** out);
SPS_API sps_status sps_batch_destroy(sps_batch* b);
/* ndt, the stream (worker) a beam runs on and its FDMT iteration count; any
 * output may be NULL */
SPS_API sps_status sps_batch_get_beam_info(const sps_batch* b, int64_t beam, int64_t* h_ndt, int32_t* h_worker,
                                           int64_t* h_niters);
/* staged passes of one beam (see sps_fdmt_set_staged); drops the graph */
SPS_API sps_status sps_batch_set_staged(sps_batch* b, int64_t beam, int staged);
SPS_API sps_status sps_batch_set_group_shape(sps_batch* b, int64_t beam, int lds_floats, int max_rows,
                                           int tiles_per_block);
/* merge schedule of one beam (see sps_fdmt_set_schedule); drops the graph */
SPS_API sps_status sps_batch_set_schedule(sps_batch* b, int64_t beam, const int32_t* h_depths, int64_t n);
SPS_API sps_status sps_batch_get_schedule(const sps_batch* b, int64_t beam, int32_t* h_depths, int64_t cap,
                                          int64_t* h_n);
/* Enqueues every beam directly (fork from / join to stream). */
SPS_API sps_status sps_batch_run(sps_batch* b, void* stream);
/* Replays the captured graph on stream (captured on the first call or after
 * a schedule change). */
SPS_API sps_status sps_batch_launch(sps_batch* b, void* stream);
/* Sorts one beam's candidates by (mdt, mt, mwidth) and returns its count
 * (synchronises stream); SPS_ERR_CAPACITY (count still set) when the beam
 * found more candidates than its cand_capacity. */
SPS_API sps_status sps_batch_sort_candidates(sps_batch* b, int64_t beam, int64_t* h_count, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SPS_H */
