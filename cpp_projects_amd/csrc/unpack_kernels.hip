// unpack_kernels.hip -- 8-bit raw-voltage unpack into overlapping complex FFT
// blocks, the input side of the coherent front-end.
//
// Both layouts produce out[pol][fft][chan][bin] (complex float, re/im
// interleaved) for pol < npol / 2, fft < nfft, chan < nchan, bin < nbin, with
// sample isamp = bin + (nbin - 2 * noverlap) * fft - noverlap of the chunk
// (0 + 0i where isamp lies outside [0, nsamp)):
//   GUPPI (CDMT/Session_guppi_gpu.cu:116-140, unpack_guppi/kernel.cu:161-187):
//     in[chan][isamp][npol], pol p's (re, im) at components 2p, 2p + 1;
//   LOFAR (CDMT/Session_lofar_gpu.cu:144-175, unpack_guppi/kernel.cu:81-122):
//     in[component][isamp][chan], pol p's (re, im) in components 2p, 2p + 1.
// Samples are signed bytes, promoted exactly to float.  The work is pure
// byte-to-float streaming (each input byte read once per block it falls in,
// 8 output bytes per complex value), bounded by HBM bandwidth.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "kernels.hpp"

namespace sps {
namespace {

// GUPPI: one lane per (chan, fft, bin); a lane reads the npol bytes of its
// sample (one 4-byte load for two polarisations; consecutive lanes,
// consecutive samples) and writes one complex per polarisation (consecutive
// lanes, consecutive bins).
__global__ void __launch_bounds__(kThreads) k_unpack_guppi(const int8_t* __restrict__ in, int64_t nsamp,
                                                           int noverlap, int nbin, int nfft, int npol, int nchan,
                                                           float2* __restrict__ out) {
    const int64_t ibin = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (ibin >= nbin) return;
    const int64_t ifft = blockIdx.y;
    const int64_t ichan = blockIdx.z;
    const int64_t isamp = ibin + (int64_t)(nbin - 2 * noverlap) * ifft - noverlap;
    const bool inside = isamp >= 0 && isamp < nsamp;
    const int8_t* src = in + (ichan * nsamp + (inside ? isamp : 0)) * npol;
    const int64_t plane = (int64_t)nfft * nchan * nbin;
    float2* dst = out + (ifft * nchan + ichan) * nbin + ibin;
    if (npol == 4 && (reinterpret_cast<uintptr_t>(in) & 3) == 0) {
        // the sample's four bytes in one load
        uint32_t w = 0u;
        if (inside) w = *reinterpret_cast<const uint32_t*>(src);
        dst[0] = make_float2((float)(int8_t)(w & 0xffu), (float)(int8_t)((w >> 8) & 0xffu));
        dst[plane] = make_float2((float)(int8_t)((w >> 16) & 0xffu), (float)(int8_t)(w >> 24));
        return;
    }
    for (int p = 0; p < npol / 2; ++p) {
        float2 v = make_float2(0.0f, 0.0f);
        if (inside) v = make_float2((float)src[2 * p], (float)src[2 * p + 1]);
        dst[p * plane] = v;
    }
}

// LOFAR: one block per (64 bins, 64 channels, fft, pol).  The tile's bytes
// are read along channels (coalesced), converted, transposed through LDS and
// written along bins (coalesced); the padded row keeps the transposed reads
// free of bank conflicts.
constexpr int kUnpackTile = 64;

__global__ void __launch_bounds__(kThreads) k_unpack_lofar(const int8_t* __restrict__ in, int64_t nsamp,
                                                           int noverlap, int nbin, int nfft, int nchan,
                                                           float2* __restrict__ out) {
    __shared__ float2 tile[kUnpackTile][kUnpackTile + 1];  // [chan][bin]
    const int64_t bin0 = (int64_t)blockIdx.x * kUnpackTile;
    const int64_t chan0 = (int64_t)blockIdx.y * kUnpackTile;
    const int64_t ifft = blockIdx.z % (unsigned)nfft;
    const int64_t pol = blockIdx.z / (unsigned)nfft;
    const int8_t* re = in + (2 * pol) * nsamp * nchan;
    const int8_t* im = in + (2 * pol + 1) * nsamp * nchan;
    const int64_t step = nbin - 2 * noverlap;
    // four channels per item: one 4-byte load of each component when the
    // channel rows are 4-byte aligned and the four lie inside the band
    const bool words = (nchan & 3) == 0 && (reinterpret_cast<uintptr_t>(in) & 3) == 0;
    constexpr int kQuads = kUnpackTile / 4;
    for (int i = (int)threadIdx.x; i < kUnpackTile * kQuads; i += kThreads) {
        const int b = i / kQuads, c = 4 * (i % kQuads);
        const int64_t ibin = bin0 + b, ichan = chan0 + c;
        const int64_t isamp = ibin + step * ifft - noverlap;
        const bool row = ibin < nbin && isamp >= 0 && isamp < nsamp;
        const int64_t k = isamp * nchan + ichan;
        if (words && row && ichan + 4 <= nchan) {
            const uint32_t wr = *reinterpret_cast<const uint32_t*>(re + k);
            const uint32_t wi = *reinterpret_cast<const uint32_t*>(im + k);
#pragma unroll
            for (int e = 0; e < 4; ++e)
                tile[c + e][b] = make_float2((float)(int8_t)((wr >> (8 * e)) & 0xffu),
                                             (float)(int8_t)((wi >> (8 * e)) & 0xffu));
        } else {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                float2 v = make_float2(0.0f, 0.0f);
                if (row && ichan + e < nchan) v = make_float2((float)re[k + e], (float)im[k + e]);
                tile[c + e][b] = v;
            }
        }
    }
    __syncthreads();
    float2* dst = out + ((pol * nfft + ifft) * nchan) * (int64_t)nbin;
    for (int i = (int)threadIdx.x; i < kUnpackTile * kUnpackTile; i += kThreads) {
        const int c = i / kUnpackTile, b = i % kUnpackTile;
        const int64_t ibin = bin0 + b, ichan = chan0 + c;
        if (ibin < nbin && ichan < nchan) dst[ichan * nbin + ibin] = tile[c][b];
    }
}

void check_unpack(int64_t nsamp, int noverlap, int nbin, int nfft, int npol, int nchan) {
    if (nsamp < 0 || noverlap < 0 || nbin < 1 || nfft < 1 || nchan < 1)
        throw Error(SPS_ERR_INVALID_ARGUMENT, "unpack sizes must be positive (nsamp, noverlap >= 0)");
    if (nbin <= 2 * noverlap) throw Error(SPS_ERR_INVALID_ARGUMENT, "nbin must exceed twice the overlap");
    if (npol != 2 && npol != 4) throw Error(SPS_ERR_INVALID_ARGUMENT, "npol must be 2 or 4 (re/im per polarisation)");
    if (nfft > 65535 || nchan > 65535) throw Error(SPS_ERR_INVALID_ARGUMENT, "nfft and nchan must be below 65536");
}

}  // namespace

void launch_unpack_guppi(const int8_t* in, int64_t nsamp, int noverlap, int nbin, int nfft, int npol, int nchan,
                         float* out, hipStream_t st) {
    check_unpack(nsamp, noverlap, nbin, nfft, npol, nchan);
    const dim3 grid((unsigned)ceil_div((int64_t)nbin, (int64_t)kThreads), (unsigned)nfft, (unsigned)nchan);
    hipLaunchKernelGGL(k_unpack_guppi, grid, dim3(kThreads), 0, st, in, nsamp, noverlap, nbin, nfft, npol, nchan,
                       reinterpret_cast<float2*>(out));
    SPS_HIP(hipGetLastError());
}

void launch_unpack_lofar(const int8_t* in, int64_t nsamp, int noverlap, int nbin, int nfft, int npol, int nchan,
                         float* out, hipStream_t st) {
    check_unpack(nsamp, noverlap, nbin, nfft, npol, nchan);
    const int64_t z = (int64_t)nfft * (npol / 2);
    if (z > 65535) throw Error(SPS_ERR_INVALID_ARGUMENT, "nfft * polarisations must be below 65536");
    const dim3 grid((unsigned)ceil_div((int64_t)nbin, (int64_t)kUnpackTile),
                    (unsigned)ceil_div((int64_t)nchan, (int64_t)kUnpackTile), (unsigned)z);
    hipLaunchKernelGGL(k_unpack_lofar, grid, dim3(kThreads), 0, st, in, nsamp, noverlap, nbin, nfft, nchan,
                       reinterpret_cast<float2*>(out));
    SPS_HIP(hipGetLastError());
}

}  // namespace sps
