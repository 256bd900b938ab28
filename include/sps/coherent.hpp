// sps/coherent.hpp -- the coherent front-end of CChunk_gpu::process over the
// libsps C ABI (CDMT/Chunk_gpu.cu:197-209,254-322; the on-the-fly chirp of
// CChunk_fly_gpu, Chunk_fly_gpu.cu:94-164).
//
// CoherentChunk holds what CChunk_gpu's constructor builds for these steps
// (the two FFT plans of create_fft_plans and the chirp buffer); its members
// are the steps: forward() is cufftExecC2C(m_fftPlanForward) on the chunk,
// waterfall(dm) is elementWiseMult, the inverse FFT and the intensity passes
// (roll_rows_normalize_sum_kernel, transpose_unpadd_intensity, dedisperse)
// for one coherent DM, leaving the [nchan * len_sft][msamp] float waterfall
// that statistical_preprocessing::calc_fdmt_inputs and the FDMT take next.
// process() is the whole of CChunk_gpu::process on one chunk: those steps,
// then per DM the statistics, the CFdmt of the keep mask and of the waterfall
// and fdmt_normalization into the DM's rows of the output image, and the
// clustering with its log.  Device pointers; asynchronous on the stream
// except process(), which synchronises for the clustering.
#pragma once

#include <hip/hip_runtime_api.h>
#include <hip/hip_vector_types.h>

#include "../sps.h"
#include <stdexcept>
#include <string>
#include <vector>

#include "fdmt.hpp"
#include "search.hpp"

class CoherentChunk {
public:
    CoherentChunk(float Fmin, float Fmax, int npol, int nchan, int len_sft, int nbin, int nfft, int noverlap,
                  float tsamp)
        : m_Fmin(Fmin), m_Fmax(Fmax), m_len_sft(len_sft), m_tsamp(tsamp) {
        sps_detail::check(sps_coherent_create(Fmin, Fmax, npol, nchan, len_sft, nbin, nfft, noverlap, tsamp, &m_h));
        sps_detail::check(sps_coherent_get_shape(m_h, &m_rows, &m_cols, &m_spectra));
        int f = 0;
        sps_detail::check(sps_coherent_is_fused(m_h, &f));
        m_fused = f != 0;
    }
    ~CoherentChunk() {
        if (m_search) (void)sps_search_destroy(m_search);
        if (m_h) (void)sps_coherent_destroy(m_h);
    }
    CoherentChunk(const CoherentChunk&) = delete;
    CoherentChunk& operator=(const CoherentChunk&) = delete;

    // waterfall rows (m_nchan * m_len_sft), columns (get_msamp()) and the
    // complex values of one chunk
    int64_t rows() const { return m_rows; }
    int64_t msamp() const { return m_cols; }
    int64_t spectra() const { return m_spectra; }
    // waterfall() runs one fused pass per DM
    bool fused() const { return m_fused; }

    // Chunk_gpu.cu:267
    void forward(float2* pcmparrRawSignalCur, hipStream_t stream = nullptr) {
        sps_detail::check(sps_coherent_forward(m_h, reinterpret_cast<float*>(pcmparrRawSignalCur), stream));
    }
    // Chunk_gpu.cu:276-322 for one coherent DM
    void waterfall(const float2* pcmparrRawSignalCur, double dm, float* parr_wfall_disp,
                   hipStream_t stream = nullptr) {
        sps_detail::check(sps_coherent_waterfall(m_h, reinterpret_cast<const float*>(pcmparrRawSignalCur), dm,
                                                 parr_wfall_disp, stream));
    }
    // the steps one at a time
    void elementWiseMult(float2* d_arrOut, const float2* d_arrInp, double dm, hipStream_t stream = nullptr) {
        sps_detail::check(sps_coherent_chirp(m_h, reinterpret_cast<const float*>(d_arrInp), dm,
                                             reinterpret_cast<float*>(d_arrOut), stream));
    }
    void inverse(float2* d_buf, hipStream_t stream = nullptr) {
        sps_detail::check(sps_coherent_inverse(m_h, reinterpret_cast<float*>(d_buf), stream));
    }
    void detect(const float2* d_buf, double dm, float* parr_wfall_disp, hipStream_t stream = nullptr) {
        sps_detail::check(
            sps_coherent_detect(m_h, reinterpret_cast<const float*>(d_buf), dm, parr_wfall_disp, stream));
    }

    // CChunk_gpu::process (Chunk_gpu.cu:254-425) on unpacked blocks, which
    // the forward FFT overwrites with their spectra.  d_image receives
    // coh_dms.size() * len_sft rows of msamp S/N values (m_pdoutImg); then
    // clusterization_main over it (widths 1..wnd, d_bins: 3 device ints,
    // the reference's {2, 1, 2}) writes filename with the waterfall's
    // sampling time len_sft * tsamp.  The reference's power-of-two gather
    // before the clustering keeps no result and is not run.  Returns
    // clusterization_main's result (false: nothing above sigma_bound).
    bool process(float2* pcmparrRawSignalCur, const std::vector<double>& coh_dms, float sigma_bound,
                 float* d_image, int wnd, const int* d_bins, const std::string& filename = "candidates.log",
                 float t_chunkBegin = 0.0f) {
        const int rows = (int)m_rows, cols = (int)m_cols;
        if (coh_dms.empty()) throw std::invalid_argument("process needs at least one coherent DM");
        if (!m_search) {
            // calc_fdmt_inputs, the CFdmt (imaxDt = len_sft) of mask and
            // waterfall and fdmt_normalization in one search call per DM; its
            // own gather (widths 1, nothing stored) is not used
            sps_search_params sp{};
            sp.f_min = m_Fmin;
            sp.f_max = m_Fmax;
            sp.tsamp = m_tsamp;
            sp.nchans = rows;
            sp.nsamps = cols;
            sp.dt_max = m_len_sft - 1;
            sp.in_dtype = SPS_F32;
            sp.norm_mode = SPS_NORM_GLOBAL;
            sp.snr_norm = 1;
            sp.box_q = 0;
            sp.threshold = 1e30f;
            sp.cand_capacity = 0;
            sps_detail::check(sps_search_create_cfdmt(&sp, m_len_sft, &m_search));
        }
        float *wf = nullptr, *thr = nullptr, *cand = nullptr, *cnt = nullptr;
        auto release = [&] {
            for (float* p : {wf, thr, cand, cnt})
                if (p) (void)hipFree(p);
        };
        try {
            hip_ok(hipMalloc(&wf, sizeof(float) * (size_t)rows * cols));
            hip_ok(hipMalloc(&thr, sizeof(float)));
            hip_ok(hipMalloc(&cand, sizeof(sps_cand)));
            hip_ok(hipMalloc(&cnt, sizeof(uint64_t)));
            hip_ok(hipMemcpy(thr, &sigma_bound, sizeof(float), hipMemcpyHostToDevice));
            forward(pcmparrRawSignalCur);
            for (size_t idm = 0; idm < coh_dms.size(); ++idm) {
                float* out = d_image + idm * (size_t)m_len_sft * cols;
                waterfall(pcmparrRawSignalCur, coh_dms[idm], wf);
                sps_detail::check(sps_search_run(m_search, wf, out, reinterpret_cast<sps_cand*>(cand),
                                                 reinterpret_cast<uint64_t*>(cnt), nullptr));
            }
            const bool ok = clusterization::clusterization_main(d_image, m_len_sft * (int)coh_dms.size(), cols, *thr,
                                                                wnd, d_bins, filename, (float)m_len_sft * m_tsamp,
                                                                t_chunkBegin);
            release();
            return ok;
        } catch (...) {
            release();
            throw;
        }
    }

private:
    static void hip_ok(hipError_t e) {
        if (e != hipSuccess) throw std::runtime_error("HIP allocation or copy failed");
    }

    float m_Fmin, m_Fmax;
    int m_len_sft;
    float m_tsamp;
    sps_search* m_search = nullptr;
    sps_coherent* m_h = nullptr;
    int64_t m_rows = 0, m_cols = 0, m_spectra = 0;
    bool m_fused = false;
};
