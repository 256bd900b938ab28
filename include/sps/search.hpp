// sps/search.hpp -- drop-in normalisation, detection and clustering calls
// over the libsps C ABI.
//
// Mirrors CDMT/Statistical_preprocessing.cuh:6-17, CDMT/Chunk_gpu.cuh:118-119
// (fdmt_normalization), Clusterization/Candidate.cuh:4-10 (Cand),
// Clusterization/Clusterization.cuh:8-68 (clusterization::),
// Clusterization/select_delegates.cuh:4-26 (select_delegates::; the
// thrust::device_vector forms are in sps/thrust_vectors.hpp), the prefix
// sums of Clusterization/cumsum_customized.cuh:11-32 and the detectors of
// detection/kernel.cu:14-20,561-640.  All array pointers are HIP
// device pointers unless named h_; thresholds and bin metrics are references
// to device memory like the reference's (Chunk_gpu.cu:373-425).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../sps.h"
#include "fdmt.hpp"

using Cand = sps_cand;  // {int mt, mdt, mwidth; float msnr}, 16 bytes
typedef sps_detection structOutDetection;

namespace statistical_preprocessing {
// Statistical_preprocessing.cuh:9-10: d_parr_inp (rows x cols) is cleaned and
// normalised in place; d_parrNormInput receives the 0/1 mask; d_buff holds at
// least 2*NRows+4 floats.
inline void calc_fdmt_inputs(float* d_parr_inp, const int NRows, const int NCols, float* d_buff,
                             float* d_parrNormInput) {
    sps_detail::check(sps_calc_fdmt_inputs(d_parr_inp, NRows, NCols, d_buff, d_parrNormInput, SPS_NORM_GLOBAL, nullptr));
}
// per-channel mean/std variant (same row cleaning)
inline void calc_fdmt_inputs_per_channel(float* d_parr_inp, const int NRows, const int NCols, float* d_buff,
                                         float* d_parrNormInput) {
    sps_detail::check(
        sps_calc_fdmt_inputs(d_parr_inp, NRows, NCols, d_buff, d_parrNormInput, SPS_NORM_PER_CHANNEL, nullptr));
}
// Statistical_preprocessing.cuh:6-7: rows whose variance lies more than
// 4 sigma from the mean row variance are zeroed in place; d_pbarrNorm (NRows
// x NCols bools, may be null) receives the keep mask; d_buff (>= NRows + 2
// floats) the row variances, their mean and standard deviation.
inline void clean_fdmtInput_and_create_normInput(float* d_arr, const int NRows, const int NCols, bool* d_pbarrNorm,
                                                 float* d_buff) {
    static_assert(sizeof(bool) == 1, "the mask is one byte per sample");
    sps_detail::check(sps_clean_fdmt_input(d_arr, NRows, NCols, reinterpret_cast<uint8_t*>(d_pbarrNorm), d_buff,
                                           nullptr));
}
inline int calcThreadsForMean_and_Disp(unsigned const int nCols) {
    int k = 0;
    while ((1u << k) < nCols && k < 10) ++k;
    return 1 << k;
}
}  // namespace statistical_preprocessing

// Chunk_gpu.cuh:118-119 (a host call here rather than a kernel launch)
// The division with a causal norm image given for its first norm_cols
// columns and constant along each row beyond (sps_fdmt_normalization_rows).
inline void fdmt_normalization_rows(const float* d_arr, int rows, int cols, const float* d_norm, int norm_cols,
                                    float* d_out, hipStream_t stream = nullptr) {
    sps_detail::check(sps_fdmt_normalization_rows(d_arr, rows, cols, d_norm, norm_cols, d_out, stream));
}

inline void fdmt_normalization(fdmt_type_* d_arr, fdmt_type_* d_norm, const int len, float* d_pOutArray) {
    sps_detail::check(sps_fdmt_normalization(d_arr, d_norm, len, d_pOutArray, nullptr));
}

namespace select_delegates {
// select_delegates.cu:16-51 over plain device arrays: group g is positions
// [d_group_begin[g], d_group_begin[g+1]) of d_indices (the last ends at
// n_indices); d_delegates[g] = the index there with the largest msnr, the
// first position on ties (thrust::max_element's rule).
inline void select_delegates(const Cand* d_cands, const int* d_indices, const int n_indices, const int* d_group_begin,
                             const int n_groups, int* d_delegates) {
    sps_detail::check(sps_select_delegates(d_cands, d_indices, n_indices, d_group_begin, n_groups, d_delegates, nullptr));
    if (hipDeviceSynchronize() != hipSuccess) throw std::runtime_error("select_delegates failed");
}
// select_delegates_kernel (select_delegates.cu:62-125), launched by the
// reference with one block per group; here a host call whose extra last
// argument is that group count (the reference's gridDim.x).
inline void select_delegates_kernel(const Cand* d_arrCandHeap, int* d_arrIndices, const int QUantCand,
                                    int* d_arrGroupBeginIndeces, int* d_arrNumDelegates, const int QUantGroups) {
    sps_detail::check(sps_select_delegates(d_arrCandHeap, d_arrIndices, QUantCand, d_arrGroupBeginIndeces, QUantGroups,
                                           d_arrNumDelegates, nullptr));
}
// select_delegates.cu:127-177 over plain device arrays: delegate i is
// d_delegates[i] and its num column the size of group i.
inline void write_log(const Cand* d_parrCand, const int QUantCandidates, const int* d_group_begin, const int n_groups,
                      const int n_indices, const int* d_delegates, const std::string& filename, const float delt_bin,
                      const float t_chunkBegin) {
    std::vector<int> begin((size_t)n_groups), dele((size_t)n_groups);
    if (n_groups > 0 &&
        (hipMemcpy(begin.data(), d_group_begin, sizeof(int) * begin.size(), hipMemcpyDeviceToHost) != hipSuccess ||
         hipMemcpy(dele.data(), d_delegates, sizeof(int) * dele.size(), hipMemcpyDeviceToHost) != hipSuccess))
        throw std::runtime_error("copy of the groups failed");
    std::vector<int64_t> num((size_t)n_groups);
    for (int g = 0; g < n_groups; ++g) num[(size_t)g] = (g + 1 < n_groups ? begin[(size_t)g + 1] : n_indices) - begin[(size_t)g];
    sps_detail::check(sps_write_log(d_parrCand, QUantCandidates, dele.data(), num.data(), n_groups, filename.c_str(),
                                    delt_bin, t_chunkBegin));
}
// the same log for delegates and group sizes already on the host
inline void write_log(const Cand* d_parrCand, const int QUantCandidates, const std::vector<int>& h_delegates,
                      const std::vector<int64_t>& h_group_sizes, const std::string& filename, const float delt_bin,
                      const float t_chunkBegin) {
    if (h_delegates.size() != h_group_sizes.size()) throw std::invalid_argument("one group size per delegate");
    sps_detail::check(sps_write_log(d_parrCand, QUantCandidates, h_delegates.data(), h_group_sizes.data(),
                                    (int64_t)h_delegates.size(), filename.c_str(), delt_bin, t_chunkBegin));
}
}  // namespace select_delegates

namespace clusterization {

// Clusterization.cuh:34-35 (__global__ digitize_kernel, launched <<<1,1024>>>
// by the reference harness): the same in-place floorf over LEnarr floats,
// as a host call on the given stream.
inline void digitize_kernel(float* d_arrfdmt, const int LEnarr, hipStream_t stream = nullptr) {
    sps_detail::check(sps_digitize(d_arrfdmt, LEnarr, stream));
}

// Clusterization.cuh:22-23: widths 1,2,..,2^QUantPower2_Wnd, mwidth = log2(w).
// *d_arrCand is hipMalloc'ed here and owned by the caller (hipFree).
inline void gather_candidates_in_dynamicalArray(const float* d_arr_fdmt, const int Rows, const int Cols,
                                                const float& val_thresh, const int QUantPower2_Wnd, Cand** d_arrCand,
                                                unsigned int* h_pquantCand) {
    int64_t n = 0;
    sps_status s = sps_boxcar_candidates(d_arr_fdmt, Rows, Cols, &val_thresh, QUantPower2_Wnd, 0, nullptr, 0, &n, nullptr);
    if (s != SPS_OK && s != SPS_ERR_CAPACITY) sps_detail::check(s);
    *h_pquantCand = (unsigned int)n;
    if (hipMalloc((void**)d_arrCand, sizeof(Cand) * (size_t)(n > 0 ? n : 1)) != hipSuccess)
        throw std::runtime_error("hipMalloc failed");
    if (n > 0)
        sps_detail::check(sps_boxcar_candidates(d_arr_fdmt, Rows, Cols, &val_thresh, QUantPower2_Wnd, 0, *d_arrCand, n,
                                                &n, nullptr));
}

// Clusterization.cuh:38-40 (Clusterization.cu:218-263), a host call here (the
// reference launches the kernel with Rows = gridDim.y): widths 1,2,..,
// 2^QUantPower2_Wnd, mwidth = log2(w), into the caller's array of
// *d_pimax_candidates_per_chunk entries (device memory); d_quantCand (device)
// receives the number found.  The reference's atomicInc wraps past the
// capacity and overwrites earlier entries; here the array holds the first
// capacity candidates in (mdt, mt, mwidth) order and the count may exceed the
// capacity.  Synchronises.
inline void gather_candidates_in_fixedArray_kernel_v0(const float* d_arr_fdmt, const int Rows, const int Cols,
                                                      const float& val_thresh, const int QUantPower2_Wnd,
                                                      unsigned int* d_pimax_candidates_per_chunk, Cand* d_arrCand,
                                                      unsigned int& d_quantCand) {
    unsigned int cap = 0;
    if (hipMemcpy(&cap, d_pimax_candidates_per_chunk, sizeof(cap), hipMemcpyDeviceToHost) != hipSuccess)
        throw std::runtime_error("hipMemcpy failed");
    int64_t n = 0;
    sps_status s = sps_boxcar_candidates(d_arr_fdmt, Rows, Cols, &val_thresh, QUantPower2_Wnd, 0, d_arrCand, cap, &n,
                                         nullptr);
    if (s != SPS_OK && s != SPS_ERR_CAPACITY) sps_detail::check(s);
    const unsigned int n32 = (unsigned int)n;
    if (hipMemcpy(&d_quantCand, &n32, sizeof(n32), hipMemcpyHostToDevice) != hipSuccess)
        throw std::runtime_error("hipMemcpy failed");
}

// Clusterization/cumsum_customized.cu:147-179,259-305: inclusive prefix sums
// (the buffer and block-size arguments of the reference's two-pass forms are
// not needed; the scratch is the library's).  Synchronise.
inline void fnc_cumsum_customized_gpu(int* d_arrInput, int* d_arr_buff, const int LEnarr, int* d_arrOutput) {
    (void)d_arr_buff;
    sps_detail::check(sps_scan_inclusive_i32(d_arrInput, d_arrOutput, LEnarr, nullptr));
}
inline void calc_cumsum_custom_gpu(int* d_arrInput, int* d_arr_buff, const int LEnarr, const int treads_per_block,
                                   int* d_arrOutput) {
    (void)d_arr_buff;
    (void)treads_per_block;
    sps_detail::check(sps_scan_inclusive_i32(d_arrInput, d_arrOutput, LEnarr, nullptr));
}
inline void inclusiveScan_CUB(float* d_arr, const int LEn) {
    sps_detail::check(sps_scan_inclusive_f32(d_arr, d_arr, LEn, nullptr));
}
inline void inclusiveScan_CUB(int* d_arr, const int LEn) {
    sps_detail::check(sps_scan_inclusive_i32(d_arr, d_arr, LEn, nullptr));
}

// Clusterization.cuh:56-63: widths 1..WndWidth, mwidth = w; false when none.
inline bool gather_candidates_in_heap_with_plan(const float* d_fdmt, const int Rows, const int Cols,
                                                const float& d_VAlTresh, const int WndWidth, Cand** d_pparrCand,
                                                unsigned int* h_piQuantCand) {
    int64_t n = 0;
    sps_status s = sps_boxcar_candidates(d_fdmt, Rows, Cols, &d_VAlTresh, WndWidth, 1, nullptr, 0, &n, nullptr);
    if (s != SPS_OK && s != SPS_ERR_CAPACITY) sps_detail::check(s);
    *h_piQuantCand = (unsigned int)n;
    if (n == 0) return false;
    if (hipMalloc((void**)d_pparrCand, sizeof(Cand) * (size_t)n) != hipSuccess) throw std::runtime_error("hipMalloc failed");
    sps_detail::check(sps_boxcar_candidates(d_fdmt, Rows, Cols, &d_VAlTresh, WndWidth, 1, *d_pparrCand, n, &n, nullptr));
    return true;
}

// Three-key gap clustering of the candidates with d_pbin_metrics and
// delegate selection: the delegates (sorted by mt) and their group sizes.
inline std::vector<int> cluster_delegates(const Cand* d_parrCand, unsigned int quant_candidates,
                                          const int* d_pbin_metrics, std::vector<int64_t>* group_sizes = nullptr) {
    std::vector<int> dele;
    if (group_sizes) group_sizes->clear();
    if (quant_candidates == 0) return dele;
    int32_t* d_del = nullptr;
    int64_t* d_gs = nullptr;
    if (hipMalloc((void**)&d_del, sizeof(int32_t) * quant_candidates) != hipSuccess ||
        hipMalloc((void**)&d_gs, sizeof(int64_t) * quant_candidates) != hipSuccess) {
        (void)hipFree(d_del);
        throw std::runtime_error("hipMalloc failed");
    }
    int64_t ng = 0;
    sps_status s = sps_cluster(d_parrCand, quant_candidates, d_pbin_metrics, d_del, d_gs, &ng, nullptr);
    std::vector<int64_t> gs((size_t)ng);
    dele.resize((size_t)ng);
    if (s == SPS_OK && ng > 0) {
        (void)hipMemcpy(dele.data(), d_del, sizeof(int32_t) * (size_t)ng, hipMemcpyDeviceToHost);
        (void)hipMemcpy(gs.data(), d_gs, sizeof(int64_t) * (size_t)ng, hipMemcpyDeviceToHost);
    }
    (void)hipFree(d_del);
    (void)hipFree(d_gs);
    sps_detail::check(s);
    if (group_sizes) *group_sizes = std::move(gs);
    return dele;
}

// Clusterization.cuh:25-32: clustering, delegates and (bwrite) the log.
inline void complete_clusterization(const Cand* d_parrCand, unsigned int quant_candidates, const int* d_pbin_metrics,
                                    const std::string& filename, const float delt_bin, const float t_chunkBegin,
                                    bool bwrite) {
    std::vector<int64_t> gs;
    std::vector<int> dele = cluster_delegates(d_parrCand, quant_candidates, d_pbin_metrics, &gs);
    if (bwrite)
        select_delegates::write_log(d_parrCand, (int)quant_candidates, dele, gs, filename, delt_bin, t_chunkBegin);
}

// Clusterization.cuh:10-20 (returns true on success; the reference's success
// path has no return statement).
inline bool clusterization_main(const float* d_fdmt, const int Rows, const int Cols, const float& d_VAlTresh,
                                const int WndWidth, const int* d_pbin_metrics, const std::string& filename,
                                const float delt_bin, const float t_chunkBegin, bool bwrite = true) {
    unsigned int n = 0;
    Cand* d_c = nullptr;
    if (!gather_candidates_in_heap_with_plan(d_fdmt, Rows, Cols, d_VAlTresh, WndWidth, &d_c, &n)) return false;
    try {
        complete_clusterization(d_c, n, d_pbin_metrics, filename, delt_bin, t_chunkBegin, bwrite);
    } catch (...) {
        (void)hipFree(d_c);
        throw;
    }
    (void)hipFree(d_c);
    return true;
}

// Clusterization.cu:310-358
inline void writeCandHostArrayToCSV(const Cand* h_arrCand, int h_quantCand, const std::string& filename) {
    std::ofstream file(filename);
    if (!file.is_open()) throw std::runtime_error("Failed to open file: " + filename);
    file << "mt,mdt,mwidth,msnr\n";
    for (int i = 0; i < h_quantCand; ++i)
        file << h_arrCand[i].mt << ',' << h_arrCand[i].mdt << ',' << h_arrCand[i].mwidth << ',' << h_arrCand[i].msnr
             << '\n';
}
inline void writeCandDeviceArrayToCSV(const Cand* d_arrCand, int h_quantCand, const std::string& filename) {
    std::vector<Cand> h((size_t)h_quantCand);
    if (h_quantCand > 0 && hipMemcpy(h.data(), d_arrCand, sizeof(Cand) * h.size(), hipMemcpyDeviceToHost) != hipSuccess)
        throw std::runtime_error("copy of candidates failed");
    writeCandHostArrayToCSV(h.data(), h_quantCand, filename);
}

// Clusterization.cu:268-308: every candidate, columns N, t bin, t sec, dedisp, width, SNR
inline void write_log_v0(const Cand* d_arrCand, const int QUantCand, const std::string& filename, const float delt_bin,
                         const float t_chunkBegin) {
    std::vector<Cand> h((size_t)QUantCand);
    if (QUantCand > 0 && hipMemcpy(h.data(), d_arrCand, sizeof(Cand) * h.size(), hipMemcpyDeviceToHost) != hipSuccess)
        throw std::runtime_error("copy of candidates failed");
    std::ofstream f(filename);
    f << std::setw(10) << "N" << std::setw(10) << "t, bin" << std::setw(10) << "t, sec" << std::setw(10) << "dedisp"
      << std::setw(10) << "width" << std::setw(10) << "SNR" << std::endl;
    for (int i = 0; i < QUantCand; ++i)
        f << std::setw(10) << i + 1 << std::setw(10) << h[(size_t)i].mt << std::setw(10)
          << t_chunkBegin + h[(size_t)i].mt * delt_bin << std::setw(10) << h[(size_t)i].mdt << std::setw(10)
          << h[(size_t)i].mwidth << std::setw(10) << h[(size_t)i].msnr << std::endl;
}

// Clusterization.cu:362-415: FREDDA candidate files (snr sampno secs boxcar idt dm beamno)
inline int countCandidatesInFreddaFile(const std::string& filename) {
    std::ifstream file(filename);
    if (!file.is_open()) return 0;
    std::string line;
    int count = 0;
    std::getline(file, line);
    while (std::getline(file, line)) ++count;
    return count;
}
inline void readFreddaCandFile(const std::string& filename, Cand* arrCand, size_t numLines) {
    std::ifstream file(filename);
    if (!file.is_open()) return;
    std::string line;
    std::getline(file, line);
    size_t index = 0;
    while (std::getline(file, line) && index < numLines) {
        std::istringstream iss(line);
        float snr, dm, secs;
        int sampno, boxcar, idt, beamno;
        if (iss >> snr >> sampno >> secs >> boxcar >> idt >> dm >> beamno) {
            arrCand[index].mt = sampno;
            arrCand[index].mdt = idt;
            arrCand[index].mwidth = boxcar;
            arrCand[index].msnr = snr;
            ++index;
        }
    }
}

}  // namespace clusterization

// The single-maximum detectors of detection/kernel.cu:561-640 with the
// reference's signatures: gridSize = (ceil(Cols / blockSize.x), Rows) and
// blockSize.x is the tile over which the window prefix sums run; the three
// auxiliary arrays hold one entry per tile (the per-tile best value, its
// row*Cols + column and its width); pstructOut (device or managed) receives
// the global best.  sz is the reference's shared-memory size and is not
// needed.  Returns pstructOut->snr >= valTresh.
namespace sps_detail {
inline bool detect(int version, fdmt_type_* d_arr, fdmt_type_* d_norm, const int Rows, const int Cols,
                   const int WndWidth, const float valTresh, const dim3 gridSize, const dim3 blockSize,
                   float* d_pAuxArray, int* d_pAuxNumArray, int* d_pWidthArray, structOutDetection* pstructOut) {
    const int tile = (int)blockSize.x;
    if (tile < 1 || (int)gridSize.x != (Cols + tile - 1) / tile || (int)gridSize.y != Rows)
        throw std::invalid_argument("gridSize must be (ceil(Cols / blockSize.x), Rows)");
    check(sps_detect_signal(version, d_arr, d_norm, Rows, Cols, WndWidth, tile, d_pAuxArray, d_pAuxNumArray,
                            d_pWidthArray, (int64_t)gridSize.x * gridSize.y, pstructOut, nullptr));
    structOutDetection h{};
    if (hipMemcpy(&h, pstructOut, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess)
        throw std::runtime_error("copy of the detection failed");
    return h.snr >= valTresh;
}
}  // namespace sps_detail

// v1: boxcar S/N of d_arr / sqrtf(1e-7 + d_norm), widths 1..WndWidth, inside a tile
inline bool doDetectSignal_v1(fdmt_type_* d_arr, fdmt_type_* d_norm, const int Rows, const int Cols,
                              const int WndWidth, const float valTresh, const dim3 gridSize, const dim3 blockSize,
                              size_t sz, float* d_pAuxArray, int* d_pAuxNumArray, int* d_pWidthArray,
                              structOutDetection* pstructOut) {
    (void)sz;
    return sps_detail::detect(1, d_arr, d_norm, Rows, Cols, WndWidth, valTresh, gridSize, blockSize, d_pAuxArray,
                              d_pAuxNumArray, d_pWidthArray, pstructOut);
}
// v2: energy ratio sqrtf(sum x^2 / sum n^2), widths 1..WndWidth
inline bool doDetectSignal_v2(fdmt_type_* d_arr, fdmt_type_* d_norm, const int Rows, const int Cols,
                              const int WndWidth, const float valTresh, const dim3 gridSize, const dim3 blockSize,
                              size_t sz, float* d_pAuxArray, int* d_pAuxNumArray, int* d_pWidthArray,
                              structOutDetection* pstructOut) {
    (void)sz;
    return sps_detail::detect(2, d_arr, d_norm, Rows, Cols, WndWidth, valTresh, gridSize, blockSize, d_pAuxArray,
                              d_pAuxNumArray, d_pWidthArray, pstructOut);
}
// v2_2: the energy ratio with a one-block final reduction (its window kernel
// is declared in a header the reference does not ship); arrBuff is not needed
inline bool doDetectSignal_v2_2(fdmt_type_* d_arr, fdmt_type_* d_norm, const int Rows, const int Cols,
                                const int WndWidth, const float valTresh, const dim3 gridSize, const dim3 blockSize,
                                size_t sz, float* d_pAuxArray, int* d_pAuxNumArray, int* d_pWidthArray,
                                structOutDetection* pstructOut, char* arrBuff) {
    (void)sz;
    (void)arrBuff;
    return sps_detail::detect(2, d_arr, d_norm, Rows, Cols, WndWidth, valTresh, gridSize, blockSize, d_pAuxArray,
                              d_pAuxNumArray, d_pWidthArray, pstructOut);
}

// The intended semantics of the detector: the global maximum over every
// window of 1..WndWidth samples inside the row (not restricted to tiles);
// ties to the smaller width, then row, then column.
inline bool doDetectSignal(fdmt_type_* d_arr, fdmt_type_* d_norm, const int Rows, const int Cols, const int WndWidth,
                           const float thr, structOutDetection* d_out) {
    sps_detail::check(sps_detect_max(d_arr, d_norm, Rows, Cols, WndWidth, d_out, nullptr));
    structOutDetection h{};
    if (hipMemcpy(&h, d_out, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess) throw std::runtime_error("copy failed");
    return h.snr >= thr;
}
