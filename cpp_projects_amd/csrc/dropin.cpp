// dropin.cpp -- C ABI entry points behind the reference's free functions
// that sit beside the search pipeline: delegate selection over caller groups
// and its log (select_delegates::), row cleaning without normalisation
// (statistical_preprocessing::clean_fdmtInput_and_create_normInput), the
// tiled single-maximum detectors (doDetectSignal_v1 / _v2 / _v2_2), the
// prefix sums of cumsum_customized and the 8-bit GUPPI / LOFAR raw-voltage
// unpack of the sessions' unpack_chunk.
#include <fstream>
#include <iomanip>
#include <type_traits>
#include <vector>

#include "abi.hpp"
#include "common.hpp"
#include "kernels.hpp"

using namespace sps;

// inclusive prefix sums: d_out (may be d_in) over n values; the scratch is
// the library's and released before returning (the call synchronises)
template <typename T, typename S>
static void scan_entry(const T* d_in, T* d_out, int64_t n, hipStream_t st, S scan) {
    SPS_REQUIRE(n >= 0, "negative length");
    if (n == 0) return;
    SPS_REQUIRE(d_in && d_out, "null argument");
    if (d_out != d_in) SPS_HIP(hipMemcpyAsync(d_out, d_in, sizeof(T) * (size_t)n, hipMemcpyDeviceToDevice, st));
    DevBuf scratch;
    scratch.ensure(sizeof(T) * (size_t)scan_scratch_entries(n));
    scan(d_out, n, scratch.as<typename std::remove_const<T>::type>(), st);
    SPS_HIP(hipStreamSynchronize(st));
}

extern "C" {

SPS_API sps_status sps_select_delegates(const sps_cand* d_cands, const int32_t* d_indices, int64_t nidx,
                                        const int32_t* d_group_begin, int64_t ngroups, int32_t* d_delegates,
                                        void* stream) {
    return guarded([&] {
        SPS_REQUIRE(nidx >= 0 && ngroups >= 0, "negative size");
        if (ngroups == 0) return;
        SPS_REQUIRE(d_cands && d_indices && d_group_begin && d_delegates, "null argument");
        launch_select_delegates(reinterpret_cast<const CandD*>(d_cands), d_indices, nidx, d_group_begin, ngroups,
                                d_delegates, as_stream(stream));
    });
}

SPS_API sps_status sps_write_log(const sps_cand* d_cands, int64_t ncand, const int32_t* h_delegates,
                                 const int64_t* h_group_sizes, int64_t ndel, const char* filename, float delt_bin,
                                 float t_chunk_begin) {
    return guarded([&] {
        SPS_REQUIRE(filename && ncand >= 0 && ndel >= 0, "invalid argument");
        SPS_REQUIRE(ndel == 0 || (h_delegates && h_group_sizes && d_cands), "null argument");
        std::vector<CandD> h((size_t)ncand);
        if (ncand > 0 && ndel > 0)
            SPS_HIP(hipMemcpy(h.data(), d_cands, sizeof(CandD) * (size_t)ncand, hipMemcpyDeviceToHost));
        std::ofstream f(filename);
        if (!f.is_open()) throw Error(SPS_ERR_RUNTIME, std::string("cannot open ") + filename);
        // columns of select_delegates::write_log (select_delegates.cu:127-177)
        f << std::setw(10) << "N" << std::setw(10) << "t, bin" << std::setw(10) << "t, sec" << std::setw(10)
          << "dedisp" << std::setw(10) << "width" << std::setw(10) << "SNR" << std::setw(10) << "num" << std::endl;
        for (int64_t i = 0; i < ndel; ++i) {
            const int32_t j = h_delegates[i];
            SPS_REQUIRE(j >= 0 && j < ncand, "delegate index out of range");
            const CandD& c = h[(size_t)j];
            const float t_sec = t_chunk_begin + (float)c.mt * delt_bin;
            f << std::setw(10) << i + 1 << std::setw(10) << c.mt << std::setw(10) << t_sec << std::setw(10) << c.mdt
              << std::setw(10) << c.mwidth << std::setw(10) << c.msnr << std::setw(10) << h_group_sizes[i]
              << std::endl;
        }
        if (!f) throw Error(SPS_ERR_RUNTIME, std::string("write failed: ") + filename);
    });
}

SPS_API sps_status sps_clean_fdmt_input(float* d_arr, int64_t nrows, int64_t ncols, uint8_t* d_mask, float* d_buff,
                                        void* stream) {
    return guarded([&] {
        SPS_REQUIRE(d_arr && nrows > 0 && ncols > 0, "invalid input");
        hipStream_t st = as_stream(stream);
        DevBuf rm, rv, stats, mu, sg, keep;
        rm.ensure(sizeof(double) * (size_t)nrows);
        rv.ensure(sizeof(double) * (size_t)nrows);
        stats.ensure(sizeof(double) * 8);
        mu.ensure(sizeof(float) * (size_t)nrows);
        sg.ensure(sizeof(float) * (size_t)nrows);
        keep.ensure((size_t)nrows);
        launch_row_stats(d_arr, 0, nrows, ncols, rm.as<double>(), rv.as<double>(), st);
        launch_stats_finalize(rm.as<double>(), rv.as<double>(), nrows, SPS_NORM_GLOBAL, stats.as<double>(),
                              mu.as<float>(), sg.as<float>(), keep.as<uint8_t>(), nullptr, st);
        launch_clean_rows(d_arr, d_mask, nrows, ncols, keep.as<uint8_t>(), rv.as<double>(), stats.as<double>(), d_buff,
                          st);
        SPS_HIP(hipStreamSynchronize(st));
    });
}

SPS_API sps_status sps_detect_signal(int version, const float* d_arr, const float* d_norm, int64_t rows, int64_t cols,
                                     int wnd, int tile, float* d_tile_val, int32_t* d_tile_pos, int32_t* d_tile_w,
                                     int64_t ntiles_cap, sps_detection* d_out, void* stream) {
    return guarded([&] {
        SPS_REQUIRE(version == 1 || version == 2, "detector version must be 1 or 2");
        SPS_REQUIRE(d_arr && d_out && d_tile_val && d_tile_pos && d_tile_w, "null argument");
        SPS_REQUIRE(rows > 0 && cols > 0 && tile > 0 && wnd > 0, "invalid sizes");
        SPS_REQUIRE(ceil_div(cols, tile) * rows <= ntiles_cap, "per-tile arrays hold fewer than rows * ceil(cols/tile)");
        launch_detect_tiles(version, d_arr, d_norm, rows, cols, tile, wnd, d_tile_val, d_tile_pos, d_tile_w, d_out,
                            as_stream(stream));
    });
}

SPS_API sps_status sps_scan_inclusive_i32(const int32_t* d_in, int32_t* d_out, int64_t n, void* stream) {
    return guarded([&] {
        // two's-complement sums are the uint32 sums of the same bits
        scan_entry<uint32_t>(reinterpret_cast<const uint32_t*>(d_in), reinterpret_cast<uint32_t*>(d_out), n,
                             as_stream(stream), scan_inclusive_u32);
    });
}

SPS_API sps_status sps_scan_inclusive_f32(const float* d_in, float* d_out, int64_t n, void* stream) {
    return guarded([&] { scan_entry<float>(d_in, d_out, n, as_stream(stream), scan_inclusive_f32); });
}

SPS_API sps_status sps_unpack_guppi(const int8_t* d_in, int64_t nsamp, int noverlap, int nbin, int nfft, int npol,
                                    int nchan, float* d_out, void* stream) {
    return guarded([&] {
        SPS_REQUIRE(d_in && d_out, "null argument");
        launch_unpack_guppi(d_in, nsamp, noverlap, nbin, nfft, npol, nchan, d_out, as_stream(stream));
    });
}

SPS_API sps_status sps_unpack_lofar(const int8_t* d_in, int64_t nsamp, int noverlap, int nbin, int nfft, int npol,
                                    int nchan, float* d_out, void* stream) {
    return guarded([&] {
        SPS_REQUIRE(d_in && d_out, "null argument");
        launch_unpack_lofar(d_in, nsamp, noverlap, nbin, nfft, npol, nchan, d_out, as_stream(stream));
    });
}

}  // extern "C"
