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
// Device pointers; asynchronous on the stream.
#pragma once

#include <hip/hip_runtime_api.h>
#include <hip/hip_vector_types.h>

#include "../sps.h"
#include "fdmt.hpp"

class CoherentChunk {
public:
    CoherentChunk(float Fmin, float Fmax, int npol, int nchan, int len_sft, int nbin, int nfft, int noverlap,
                  float tsamp) {
        sps_detail::check(sps_coherent_create(Fmin, Fmax, npol, nchan, len_sft, nbin, nfft, noverlap, tsamp, &m_h));
        sps_detail::check(sps_coherent_get_shape(m_h, &m_rows, &m_cols, &m_spectra));
        int f = 0;
        sps_detail::check(sps_coherent_is_fused(m_h, &f));
        m_fused = f != 0;
    }
    ~CoherentChunk() {
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

private:
    sps_coherent* m_h = nullptr;
    int64_t m_rows = 0, m_cols = 0, m_spectra = 0;
    bool m_fused = false;
};
