// coherent_kernels.hip -- the coherent front-end between the 8-bit unpack and
// the incoherent search: CChunk_gpu::process with CChunk_fly_gpu's chirp
// (CDMT/Chunk_gpu.cu:254-322, Chunk_fly_gpu.cu:94-164).
//
// The chunk holds out[pol][fft][chan][bin] complex FFT blocks (the unpack's
// layout, npol/2 polarisations, nbin = len_sft * mbin bins).  Per chunk:
//  1. forward FFT of every block (nbin points, in place; rocFFT);
// per coherent DM:
//  2. chirp: each spectrum bin times the phase factor of its frequency, its
//     sub-band taper and 1/nbin, the spectrum half-rotated (fftshift), into a
//     second buffer (k_coherent_chirp);
//  3. inverse FFT of every mbin-bin sub-band (rocFFT): len_sft sub-channel
//     time series of mbin samples per block;
//  4. the intensity waterfall (k_coherent_detect, one pass): per sample the
//     polarisation sum of |x|^2 / nbin^2, the block rotated by half a row,
//     the sub-channel matrices transposed with both frequency axes reversed
//     and the overlap trimmed, and each row shifted by its incoherent
//     dispersion delay -- the three reference passes roll_rows_normalize_
//     sum_kernel, transpose_unpadd_intensity and dedisperse composed into
//     one gather, with the same arithmetic per value.
// The waterfall has nchan * len_sft rows of msamp = nfft * (mbin - 2 *
// noverlap / len_sft) samples.  All of it is HBM-bound streaming; the FFTs
// are rocFFT's.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <cmath>
#include <memory>

#include "abi.hpp"
#include "common.hpp"
#include "kernels.hpp"

namespace sps {
namespace {

struct CoherentShape {
    int npol, nchan, len_sft, nbin, nfft, noverlap;
    double fmin, fmax;
    float tsamp;
    int mbin() const { return nbin / len_sft; }
    int nov_sub() const { return noverlap / len_sft; }
    int mbin_adj() const { return mbin() - 2 * nov_sub(); }
    int64_t msamp() const { return (int64_t)nfft * mbin_adj(); }
    int64_t rows() const { return (int64_t)nchan * len_sft; }
    int64_t spectra() const { return (int64_t)nfft * nchan * (npol / 2) * nbin; }  // complex values
};

// kernel_el_wise_mult_onthe_fly (Chunk_fly_gpu.cu:119-164): one thread per
// (chan, sub-band, bin); the factor for every block of that channel.  The
// sub-band centre is taken to float as in the reference, the rest in double.
__global__ void __launch_bounds__(kThreads) k_coherent_chirp(const float2* __restrict__ in, float2* __restrict__ out,
                                                             double dm, int nchan, int len_sft, int mbin,
                                                             int nblocks, double fmin, double bw_sub,
                                                             double bw_chan) {
    const int ibin = (int)(blockIdx.x * kThreads + threadIdx.x);
    if (ibin >= mbin) return;
    const int isft = (int)blockIdx.y;
    const int ichan = (int)blockIdx.z;
    const float f0 = (float)(fmin + bw_sub * (0.5 + ichan) + bw_chan * (isft - len_sft / 2.0 + 0.5));
    const double bf = -0.5 * bw_chan + (ibin + 0.5) * bw_chan / mbin;
    const double taper = 1.0 / sqrt(1.0 + pow(bf / (0.47 * bw_chan), 80.0));
    const double r = bf / (double)f0;
    const double delay = dm * r * r / ((double)f0 + bf) * 4.148808e9;
    double whole;
    const double t = -modf(delay, &whole) * 2.0;
    double s, c;
    sincospi(t, &s, &c);
    const int nbin = mbin * len_sft;
    const int col = (isft * mbin + ibin + nbin / 2) % nbin;
    const float2 d = make_float2((float)(c * taper / (double)nbin), (float)(s * taper / (double)nbin));
    int64_t idx = (int64_t)ichan * nbin + col;
    const int64_t stride = (int64_t)nbin * nchan;
    // four blocks' loads in flight before their products are stored
    int b = 0;
    for (; b + 4 <= nblocks; b += 4, idx += 4 * stride) {
        float2 a[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) a[u] = in[idx + u * stride];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            out[idx + u * stride] = make_float2(a[u].x * d.x - a[u].y * d.y, a[u].x * d.y + a[u].y * d.x);
    }
    for (; b < nblocks; ++b, idx += stride) {
        const float2 a = in[idx];
        out[idx] = make_float2(a.x * d.x - a.y * d.y, a.x * d.y + a.y * d.x);
    }
}

// The intensity waterfall of one DM, wf[row][t] (see the file comment),
// one thread per output sample: the source sample of the dispersion shift,
// then of the transpose and trim, then of the half-row rotation, and the
// polarisation sum of ((x*x + y*y) / nbin) / nbin in double, each term
// rounded to float and added in float (roll_rows_normalize_sum_kernel,
// Chunk_gpu.cu:476-493; transpose_unpadd_intensity, :497-515; dedisperse,
// :613-632).
// inv_nbin2 > 0: nbin is a power of two and inv_nbin2 = nbin^-2 exactly, so
// the two divisions are this one exact scaling (each rounds the same real
// value once); otherwise the divisions themselves.
__global__ void __launch_bounds__(kThreads) k_coherent_detect(const float2* __restrict__ buf, float* __restrict__ wf,
                                                              int npol, int nchan, int len_sft, int mbin,
                                                              int nov_sub, int mbin_adj, int nfft, int msamp,
                                                              double dm, double fmin, double foff, double tsamp_wf,
                                                              double inv_nbin2) {
    const int t = (int)(blockIdx.x * kThreads + threadIdx.x);
    if (t >= msamp) return;
    const int row = (int)blockIdx.y;
    const int nrows = nchan * len_sft;
    // dedisperse: wf[row][(isamp + ishift) % msamp] = intensity[row][isamp]
    // (the row's shift is block-uniform; reduced modulo msamp once)
    const double fr = (double)(nrows - 1 - row) * foff + fmin;
    const double delay = 4.148808e3 * dm * (1.0 / (fmin * fmin) - 1.0 / (fr * fr));
    const int64_t ishift = (int64_t)round(delay / tsamp_wf);
    const int sh = (int)(ishift % msamp);  // dm >= 0, so ishift >= 0
    int isamp = t - sh;
    if (isamp < 0) isamp += msamp;
    // transpose_unpadd_intensity: row = isub * len_sft + len_sft - 1 - ich
    const int isub = row / len_sft;
    const int ich = len_sft - 1 - row % len_sft;
    const int ifft = isamp / mbin_adj;
    const int ibin = isamp - ifft * mbin_adj + nov_sub;
    const int c = nchan - 1 - isub;
    // roll_rows_normalize_sum_kernel: rolled[j] = f(block[(j - nbin/2) mod nbin])
    const int nbin = mbin * len_sft;
    int src = ich * mbin + ibin - nbin / 2;
    if (src < 0) src += nbin;
    int64_t idx = ((int64_t)ifft * nchan + c) * nbin + src;
    const int64_t plane = (int64_t)nfft * nchan * nbin;
    float v = 0.0f;
    for (int p = 0; p < npol / 2; ++p, idx += plane) {
        const float2 a = buf[idx];
        const double x = a.x, y = a.y;
        const double e = x * x + y * y;
        v += (float)(inv_nbin2 > 0.0 ? e * inv_nbin2 : e / nbin / nbin);
    }
    wf[(int64_t)row * msamp + t] = v;
}

void check_hipfft(hipfftResult r, const char* what) {
    if (r != HIPFFT_SUCCESS) throw Error(SPS_ERR_RUNTIME, std::string("hipFFT: ") + what);
}

}  // namespace
}  // namespace sps

using namespace sps;

struct sps_coherent {
    CoherentShape sh{};
    hipfftHandle fwd = 0, inv = 0;
    bool have_fwd = false, have_inv = false;
    DevBuf chirped;  // the chirped spectra, then their sub-band time series
    ~sps_coherent() {
        if (have_fwd) (void)hipfftDestroy(fwd);
        if (have_inv) (void)hipfftDestroy(inv);
    }
};

namespace {

void coherent_chirp(const sps_coherent& h, const float* d_in, double dm, float* d_out, hipStream_t st) {
    const CoherentShape& s = h.sh;
    // double bw = m_Fmax - m_Fmin: a float difference (Chunk_fly_gpu.cu:96)
    const double bw = (double)((float)s.fmax - (float)s.fmin);
    const double bw_sub = bw / s.nchan;
    const double bw_chan = bw_sub / s.len_sft;
    const dim3 grid((unsigned)ceil_div((int64_t)s.mbin(), (int64_t)kThreads), (unsigned)s.len_sft,
                    (unsigned)s.nchan);
    hipLaunchKernelGGL(k_coherent_chirp, grid, dim3(kThreads), 0, st, reinterpret_cast<const float2*>(d_in),
                       reinterpret_cast<float2*>(d_out), dm, s.nchan, s.len_sft, s.mbin(), s.nfft * (s.npol / 2),
                       s.fmin, bw_sub, bw_chan);
    SPS_HIP(hipGetLastError());
}

void coherent_detect(const sps_coherent& h, const float* d_buf, double dm, float* d_wf, hipStream_t st) {
    const CoherentShape& s = h.sh;
    const double foff = (s.fmax - s.fmin) / ((double)s.nchan * s.len_sft);
    // (double)(m_len_sft * m_tsamp): the product is taken in float
    const double tsamp_wf = (double)((float)(unsigned)s.len_sft * s.tsamp);
    const dim3 grid((unsigned)ceil_div(s.msamp(), (int64_t)kThreads), (unsigned)s.rows());
    const bool pow2 = (s.nbin & (s.nbin - 1)) == 0;
    const double inv_nbin2 = pow2 ? std::ldexp(1.0, -2 * (int)std::log2((double)s.nbin)) : 0.0;
    hipLaunchKernelGGL(k_coherent_detect, grid, dim3(kThreads), 0, st, reinterpret_cast<const float2*>(d_buf), d_wf,
                       s.npol, s.nchan, s.len_sft, s.mbin(), s.nov_sub(), s.mbin_adj(), s.nfft, (int)s.msamp(), dm,
                       s.fmin, foff, tsamp_wf, inv_nbin2);
    SPS_HIP(hipGetLastError());
}

void coherent_fft(sps_coherent& h, hipfftHandle plan, float* d, int dir, hipStream_t st) {
    check_hipfft(hipfftSetStream(plan, st), "set stream");
    check_hipfft(hipfftExecC2C(plan, reinterpret_cast<hipfftComplex*>(d), reinterpret_cast<hipfftComplex*>(d), dir),
                 "C2C");
}

}  // namespace

extern "C" {

SPS_API sps_status sps_coherent_create(float f_min, float f_max, int npol, int nchan, int len_sft, int nbin,
                                       int nfft, int noverlap, float tsamp, sps_coherent** out) {
    return guarded([&] {
        SPS_REQUIRE(out, "null argument");
        *out = nullptr;
        SPS_REQUIRE(npol == 2 || npol == 4, "npol must be 2 or 4 (re/im per polarisation)");
        SPS_REQUIRE(nchan >= 1 && len_sft >= 1 && nfft >= 1 && noverlap >= 0, "sizes must be positive");
        SPS_REQUIRE(nbin >= len_sft && nbin % len_sft == 0, "nbin must be a multiple of len_sft");
        SPS_REQUIRE(f_max > f_min && tsamp > 0.0f, "band and sampling time must be positive");
        std::unique_ptr<sps_coherent> h(new sps_coherent());
        h->sh = CoherentShape{npol, nchan, len_sft, nbin, nfft, noverlap, (double)f_min, (double)f_max, tsamp};
        SPS_REQUIRE(h->sh.mbin_adj() >= 1, "the overlap leaves no samples per block");
        SPS_REQUIRE(h->sh.spectra() < ((int64_t)1 << 31), "chunk too large");
        SPS_REQUIRE(h->sh.rows() <= 65535, "nchan * len_sft must be below 65536");
        SPS_REQUIRE(h->sh.msamp() < ((int64_t)1 << 31), "waterfall rows too long");
        // create_fft_plans (Chunk_gpu.cu:197-209)
        int n1 = nbin, n2 = h->sh.mbin();
        check_hipfft(hipfftPlanMany(&h->fwd, 1, &n1, nullptr, 1, nbin, nullptr, 1, nbin, HIPFFT_C2C,
                                    nfft * nchan * (npol / 2)),
                     "forward plan");
        h->have_fwd = true;
        check_hipfft(hipfftPlanMany(&h->inv, 1, &n2, nullptr, 1, n2, nullptr, 1, n2, HIPFFT_C2C,
                                    nfft * nchan * len_sft * (npol / 2)),
                     "inverse plan");
        h->have_inv = true;
        h->chirped.ensure(sizeof(float) * 2 * (size_t)h->sh.spectra());
        *out = h.release();
    });
}

SPS_API sps_status sps_coherent_destroy(sps_coherent* h) {
    return guarded([&] { delete h; });
}

SPS_API sps_status sps_coherent_get_shape(const sps_coherent* h, int64_t* h_rows, int64_t* h_cols,
                                          int64_t* h_spectra) {
    return guarded([&] {
        SPS_REQUIRE(h, "null handle");
        if (h_rows) *h_rows = h->sh.rows();
        if (h_cols) *h_cols = h->sh.msamp();
        if (h_spectra) *h_spectra = h->sh.spectra();
    });
}

SPS_API sps_status sps_coherent_forward(sps_coherent* h, float* d_blocks, void* stream) {
    return guarded([&] {
        SPS_REQUIRE(h && d_blocks, "null argument");
        coherent_fft(*h, h->fwd, d_blocks, HIPFFT_FORWARD, as_stream(stream));
    });
}

SPS_API sps_status sps_coherent_chirp(sps_coherent* h, const float* d_spectra, double dm, float* d_out,
                                      void* stream) {
    return guarded([&] {
        SPS_REQUIRE(h && d_spectra && d_out, "null argument");
        SPS_REQUIRE(dm >= 0.0 && std::isfinite(dm), "the coherent DM must be finite and non-negative");
        coherent_chirp(*h, d_spectra, dm, d_out, as_stream(stream));
    });
}

SPS_API sps_status sps_coherent_inverse(sps_coherent* h, float* d_buf, void* stream) {
    return guarded([&] {
        SPS_REQUIRE(h && d_buf, "null argument");
        coherent_fft(*h, h->inv, d_buf, HIPFFT_BACKWARD, as_stream(stream));
    });
}

SPS_API sps_status sps_coherent_detect(sps_coherent* h, const float* d_buf, double dm, float* d_waterfall,
                                       void* stream) {
    return guarded([&] {
        SPS_REQUIRE(h && d_buf && d_waterfall, "null argument");
        SPS_REQUIRE(dm >= 0.0 && std::isfinite(dm), "the coherent DM must be finite and non-negative");
        coherent_detect(*h, d_buf, dm, d_waterfall, as_stream(stream));
    });
}

SPS_API sps_status sps_coherent_waterfall(sps_coherent* h, const float* d_spectra, double dm, float* d_waterfall,
                                          void* stream) {
    return guarded([&] {
        SPS_REQUIRE(h && d_spectra && d_waterfall, "null argument");
        SPS_REQUIRE(dm >= 0.0 && std::isfinite(dm), "the coherent DM must be finite and non-negative");
        hipStream_t st = as_stream(stream);
        float* buf = h->chirped.as<float>();
        coherent_chirp(*h, d_spectra, dm, buf, st);
        coherent_fft(*h, h->inv, buf, HIPFFT_BACKWARD, st);
        coherent_detect(*h, buf, dm, d_waterfall, st);
    });
}

}  // extern "C"
