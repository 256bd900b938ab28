// sps/unpack.hpp -- the sessions' 8-bit raw-voltage unpack over the libsps C
// ABI.
//
// Mirrors the kernels the coherent sessions launch from unpack_chunk:
// unpack_chunk_guppi_gpu (CDMT/Session_guppi_gpu.cu:116-140, launched at
// :53-60) and unpackInput_lofar (CDMT/Session_lofar_gpu.cu:144-175, launched
// at :54-67); the standalone versions are unpack_guppi/kernel.cu:81-122,
// 161-187.  inp_type_ is char (MyVSLib_C++/Constants.h:17): signed 8-bit
// samples.  The launches are host calls here: the kernel shape is the
// library's, and the grid and block arguments of the reference's <<<>>>
// launches have no counterpart.  d_parrOut holds nfft * nchan * nbin *
// npol/2 complex values laid out [pol][fft][chan][bin].
#pragma once

#include <hip/hip_runtime_api.h>
#include <hip/hip_vector_types.h>

#include "../sps.h"
#include "fdmt.hpp"

// GUPPI: parrInput[ichan][isamp][npol] (Session_guppi_gpu.cu:116-140)
inline void unpack_chunk_guppi_gpu(int nsamp, int Noverlap, int nbin, int nfft, int npol, int nchan,
                                   const char* parrInput, float2* d_parrOut, hipStream_t stream = nullptr) {
    sps_detail::check(sps_unpack_guppi(reinterpret_cast<const int8_t*>(parrInput), nsamp, Noverlap, nbin, nfft, npol,
                                       nchan, reinterpret_cast<float*>(d_parrOut), stream));
}

// LOFAR: d_parrInput[component][isamp][chan], NPolPh = npol / 2 complex
// polarisations (Session_lofar_gpu.cu:144-175)
inline void unpackInput_lofar(float2* pcmparrRawSignalCur, const char* d_parrInput, const int NPolPh,
                              const int LenChunk, const int nbin, const int Noverlap, const int nfft,
                              const int nchan, hipStream_t stream = nullptr) {
    sps_detail::check(sps_unpack_lofar(reinterpret_cast<const int8_t*>(d_parrInput), LenChunk, Noverlap, nbin, nfft,
                                       2 * NPolPh, nchan, reinterpret_cast<float*>(pcmparrRawSignalCur), stream));
}
