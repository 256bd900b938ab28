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

#include <algorithm>
#include <cmath>
#include <memory>
#include <vector>

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

// The sub-band offset of bin ibin and its taper 1/sqrt(1 + (bf/(0.47
// bw_chan))^80), in double (Chunk_fly_gpu.cu:139-146).  The taper depends on
// the bin alone, so the host tabulates it once per geometry.
__host__ __device__ inline double chirp_bf(int ibin, int mbin, double bw_chan) {
    return -0.5 * bw_chan + (ibin + 0.5) * bw_chan / mbin;
}
inline double chirp_taper(int ibin, int mbin, double bw_chan) {
    const double bf = chirp_bf(ibin, mbin, bw_chan);
    return 1.0 / std::sqrt(1.0 + std::pow(bf / (0.47 * bw_chan), 80.0));
}

// The chirp factor of (chan, sub-band, bin) (kernel_el_wise_mult_onthe_fly,
// Chunk_fly_gpu.cu:119-164): the sub-band centre is taken to float as in the
// reference, the rest in double; the factor includes the taper and 1/nbin
// (inv_nbin > 0: nbin is a power of two and the division is this exact
// scaling).  *col receives its column in the half-rotated spectrum row.
__device__ inline float2 chirp_factor(int ichan, int isft, int ibin, int len_sft, int mbin, double dm, double fmin,
                                      double bw_sub, double bw_chan, const double* __restrict__ taper_tab,
                                      double inv_nbin, int* col) {
    const float f0 = (float)(fmin + bw_sub * (0.5 + ichan) + bw_chan * (isft - len_sft / 2.0 + 0.5));
    const double bf = chirp_bf(ibin, mbin, bw_chan);
    const double taper = taper_tab[ibin];
    const double r = bf / (double)f0;
    const double delay = dm * r * r / ((double)f0 + bf) * 4.148808e9;
    double whole;
    const double t = -modf(delay, &whole) * 2.0;
    double s, c;
    sincospi(t, &s, &c);
    const int nbin = mbin * len_sft;
    *col = (isft * mbin + ibin + nbin / 2) % nbin;
    if (inv_nbin > 0.0) return make_float2((float)(c * taper * inv_nbin), (float)(s * taper * inv_nbin));
    return make_float2((float)(c * taper / (double)nbin), (float)(s * taper / (double)nbin));
}

// a * d, each product and sum rounded (no contraction)
__device__ inline float2 chirp_mul(float2 a, float2 d) {
    return make_float2(a.x * d.x - a.y * d.y, a.x * d.y + a.y * d.x);
}

// One thread per (chan, sub-band, bin); the factor for every block of that
// channel.
__global__ void __launch_bounds__(kThreads) k_coherent_chirp(const float2* __restrict__ in, float2* __restrict__ out,
                                                             double dm, int nchan, int len_sft, int mbin,
                                                             int nblocks, double fmin, double bw_sub,
                                                             double bw_chan, const double* __restrict__ taper_tab,
                                                             double inv_nbin) {
    const int ibin = (int)(blockIdx.x * kThreads + threadIdx.x);
    if (ibin >= mbin) return;
    const int isft = (int)blockIdx.y;
    const int ichan = (int)blockIdx.z;
    const int nbin = mbin * len_sft;
    int col;
    const float2 d = chirp_factor(ichan, isft, ibin, len_sft, mbin, dm, fmin, bw_sub, bw_chan, taper_tab, inv_nbin,
                                  &col);
    int64_t idx = (int64_t)ichan * nbin + col;
    const int64_t stride = (int64_t)nbin * nchan;
    // four blocks' loads in flight before their products are stored
    int b = 0;
    for (; b + 4 <= nblocks; b += 4, idx += 4 * stride) {
        float2 a[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) a[u] = in[idx + u * stride];
#pragma unroll
        for (int u = 0; u < 4; ++u) out[idx + u * stride] = chirp_mul(a[u], d);
    }
    for (; b < nblocks; ++b, idx += stride) out[idx] = chirp_mul(in[idx], d);
}

// The factors alone, fac[chan][col] (the spectrum rows' layout), for the
// fused pass.
__global__ void __launch_bounds__(kThreads) k_coherent_factors(float2* __restrict__ fac, double dm, int len_sft,
                                                               int mbin, double fmin, double bw_sub, double bw_chan,
                                                               const double* __restrict__ taper_tab,
                                                               double inv_nbin) {
    const int ibin = (int)(blockIdx.x * kThreads + threadIdx.x);
    if (ibin >= mbin) return;
    const int isft = (int)blockIdx.y;
    const int ichan = (int)blockIdx.z;
    int col;
    const float2 d = chirp_factor(ichan, isft, ibin, len_sft, mbin, dm, fmin, bw_sub, bw_chan, taper_tab, inv_nbin,
                                  &col);
    fac[(int64_t)ichan * mbin * len_sft + col] = d;
}

// dedisperse (Chunk_gpu.cu:613-632): the shift of waterfall row `row`,
// round(4.148808e3 * dm * (fmin^-2 - f_row^-2) / tsamp_wf), reduced modulo
// msamp (dm >= 0, so the shift is >= 0).
__device__ inline int dedisp_shift(int row, int nrows, int msamp, double dm, double fmin, double foff,
                                   double tsamp_wf) {
    const double fr = (double)(nrows - 1 - row) * foff + fmin;
    const double delay = 4.148808e3 * dm * (1.0 / (fmin * fmin) - 1.0 / (fr * fr));
    const int64_t ishift = (int64_t)round(delay / tsamp_wf);
    return (int)(ishift % msamp);
}

// ((x*x + y*y) / nbin) / nbin in double, rounded to float; inv_nbin2 > 0
// means nbin is a power of two and inv_nbin2 = nbin^-2 exactly, and the one
// exact scaling rounds the same real value once as the two divisions do.
__device__ inline float pol_intensity(float2 a, int nbin, double inv_nbin2) {
    const double x = a.x, y = a.y;
    const double e = x * x + y * y;
    return (float)(inv_nbin2 > 0.0 ? e * inv_nbin2 : e / nbin / nbin);
}

// The intensity waterfall of one DM, wf[row][t] (see the file comment),
// one thread per output sample: the source sample of the dispersion shift,
// then of the transpose and trim, then of the half-row rotation, and the
// polarisation sum of ((x*x + y*y) / nbin) / nbin in double, each term
// rounded to float and added in float (roll_rows_normalize_sum_kernel,
// Chunk_gpu.cu:476-493; transpose_unpadd_intensity, :497-515; dedisperse,
// :613-632).
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
    int isamp = t - dedisp_shift(row, nrows, msamp, dm, fmin, foff, tsamp_wf);
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
    for (int p = 0; p < npol / 2; ++p, idx += plane) v += pol_intensity(buf[idx], nbin, inv_nbin2);
    wf[(int64_t)row * msamp + t] = v;
}

// ---- the fused per-DM pass: chirp, inverse FFT and detection in one kernel
//
// One block per (channel c, sub-band segment k, run of blocks): for each block
// and polarisation it loads the segment's mbin = N spectrum values, multiplies
// them by their chirp factors, takes the unnormalised inverse DFT of length N
// (exp(+2 pi i n q / N), as the inverse FFT of step 3), and adds each output's
// intensity over the polarisations in registers; the kept samples go straight
// to their dedispersed waterfall positions.  The segment's output lands in
// one waterfall row: the half-row rotation moves segment k to sub-channel
// ich = (k + len_sft/2) mod len_sft (len_sft even, so nbin/2 is a whole number
// of segments).
//
// The DFT is a Stockham autosort FFT over N = 8^a * {1, 2, 4} points with
// T = N/8 lanes: lane t always holds elements t + T*m (m = 0..7) of the
// current sequence; a pass of radix R does butterflies j = t + T*u
// (u < 8/R) on x[j + r*N/R] = z[u + r*8/R], twiddles them by
// exp(2 pi i r (j mod Ns) / (Ns R)) and writes output q of butterfly j to
// (j / Ns) * Ns * R + (j mod Ns) + q * Ns, where Ns is the product of the
// earlier radices.  After the last pass (Ns R = N) that position is
// t + T*(u + q*8/R), so the lane keeps its outputs in registers.  Between
// passes the sequence goes through LDS with a pass-dependent padding (one
// complex every 32, then every 8, then none) that keeps the scattered writes
// and linear reads near conflict-free.

template <int R>
__device__ inline void dft_inv(float2 (&v)[R]);

__device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ inline float2 mul_i(float2 a) { return make_float2(-a.y, a.x); }  // a * i
// a * w with fused multiply-adds (the FFT's own rounding, not the reference's)
__device__ inline float2 cmul_fma(float2 a, float2 w) {
    return make_float2(fmaf(a.x, w.x, -a.y * w.y), fmaf(a.x, w.y, a.y * w.x));
}

template <>
__device__ inline void dft_inv<2>(float2 (&v)[2]) {
    const float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <>
__device__ inline void dft_inv<4>(float2 (&v)[4]) {
    const float2 s0 = cadd(v[0], v[2]), s1 = csub(v[0], v[2]);
    const float2 s2 = cadd(v[1], v[3]), s3 = mul_i(csub(v[1], v[3]));
    v[0] = cadd(s0, s2);
    v[2] = csub(s0, s2);
    v[1] = cadd(s1, s3);
    v[3] = csub(s1, s3);
}

template <>
__device__ inline void dft_inv<8>(float2 (&v)[8]) {
    float2 e[4] = {v[0], v[2], v[4], v[6]};
    float2 o[4] = {v[1], v[3], v[5], v[7]};
    dft_inv<4>(e);
    dft_inv<4>(o);
    constexpr float h = 0.70710678118654752440f;
    // o[q] * exp(2 pi i q / 8): o1 = a * h, o3 = b * h, the scaling fused into
    // the outputs' additions
    const float2 a = make_float2(o[1].x - o[1].y, o[1].x + o[1].y);
    const float2 b = make_float2(-(o[3].x + o[3].y), o[3].x - o[3].y);
    const float2 o2 = mul_i(o[2]);
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = make_float2(fmaf(a.x, h, e[1].x), fmaf(a.y, h, e[1].y));
    v[5] = make_float2(fmaf(-a.x, h, e[1].x), fmaf(-a.y, h, e[1].y));
    v[2] = cadd(e[2], o2);
    v[6] = csub(e[2], o2);
    v[3] = make_float2(fmaf(b.x, h, e[3].x), fmaf(b.y, h, e[3].y));
    v[7] = make_float2(fmaf(-b.x, h, e[3].x), fmaf(-b.y, h, e[3].y));
}

template <int SH>
__device__ inline int fft_pad(int i) {
    return SH >= 31 ? i : i + (i >> SH);
}

template <int P>
struct FftPad {
    static constexpr int sh = P == 0 ? 5 : (P == 1 ? 3 : 31);
};

// The pass structure of an N-point transform: radix 8 while at least 8
// points per butterfly remain, then 2 or 4; a pass's stride Ns is the
// product of the earlier radices.  Passes after the first read their
// twiddles exp(2 pi i r kk / (Ns R)) (r = 1..R-1, kk < Ns) from a table
// tw[off(Ns) + (r - 1) * Ns + kk] (fft_twiddle_table on the host, in double
// rounded to float) that each launch block copies into LDS.
template <int N>
struct FftPlan {
    static constexpr int T = N / 8;
    static constexpr int radix(int ns) { return (N / ns) >= 8 ? 8 : N / ns; }
    static constexpr int tw_off(int ns) {
        int o = 0;
        for (int n = 8; n < ns; n *= radix(n)) o += (radix(n) - 1) * n;
        return o;
    }
    static constexpr int tw_size = tw_off(N);
    static constexpr int data_size = N + N / 8;  // the widest padding
};

// Pass P of stride NS on the lane's elements z[m] = x[t + T*m].  With the
// paddings used (pass 0: one complex in 32, its butterflies writing 8
// consecutive slots from a multiple of 8; pass 1: one in 8 at stride 8;
// later: none) a padded index base + c never carries across a padding
// boundary, so pad(base + c) = pad(base) + c + (c >> sh): every LDS access is
// a per-lane base plus a compile-time offset.
template <int N, int NS, int P>
__device__ inline void fft_passes(float2 (&z)[8], float2* lds, const float2* tw, int t) {
    using F = FftPlan<N>;
    constexpr int T = F::T;
    constexpr int R = F::radix(NS);
    constexpr int U = 8 / R;
    constexpr bool last = NS * R == N;
    static_assert(P != 0 || (NS == 1 && R == 8), "pass 0 is radix 8");
    static_assert(P != 1 || NS == 8, "pass 1 has stride 8");
    float2 o[8];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        float2 v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = z[u + r * U];
        if constexpr (NS > 1) {
            const float2* w = tw + F::tw_off(NS) + ((t + T * u) & (NS - 1));
#pragma unroll
            for (int r = 1; r < R; ++r) v[r] = cmul_fma(v[r], w[(r - 1) * NS]);
        }
        dft_inv<R>(v);
#pragma unroll
        for (int q = 0; q < R; ++q) o[u + q * U] = v[q];
    }
    if constexpr (last) {
#pragma unroll
        for (int m = 0; m < 8; ++m) z[m] = o[m];
    } else {
        constexpr int sh = FftPad<P>::sh;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = t + T * u;
            const int kk = j & (NS - 1);
            float2* wp = lds + fft_pad<sh>((j - kk) * R + kk);
#pragma unroll
            for (int q = 0; q < R; ++q) wp[q * NS + ((q * NS) >> sh)] = o[u + q * U];
        }
        __syncthreads();
        const float2* rp = lds + fft_pad<sh>(t);
#pragma unroll
        for (int m = 0; m < 8; ++m) z[m] = rp[T * m + ((T * m) >> sh)];
        __syncthreads();
        fft_passes<N, NS * R, P + 1>(z, lds, tw, t);
    }
}

// What chirp_factor needs besides the bin: the sub-band and channel widths,
// the taper table and the exact 1/nbin (0 when nbin is not a power of two).
struct ChirpArgs {
    double bw_sub, bw_chan;
    const double* taper;
    double inv_nbin;
};

// fac == nullptr: each launch block computes its segment's factors itself
// (the same chirp_factor values the table would hold)
template <int N>
__global__ void __launch_bounds__(N / 8) k_coherent_fused(const float2* __restrict__ spec,
                                                          const float2* __restrict__ fac, ChirpArgs ca,
                                                          const float2* __restrict__ twg, float* __restrict__ wf,
                                                          int npairs, int nchan, int len_sft, int nfft, int bpw,
                                                          int nov_sub, int mbin_adj, int msamp, double dm,
                                                          double fmin, double foff, double tsamp_wf, float scale) {
    using F = FftPlan<N>;
    constexpr int T = F::T;
    __shared__ float2 lds[F::data_size];
    __shared__ float2 tw[F::tw_size];
    const int t = (int)threadIdx.x;
    for (int e = t; e < F::tw_size; e += T) tw[e] = twg[e];
    const int seg = (int)blockIdx.x;
    const int c = seg / len_sft;
    const int k = seg - c * len_sft;
    const int ich = (k + len_sft / 2) % len_sft;
    const int nrows = nchan * len_sft;
    const int row = (nchan - 1 - c) * len_sft + len_sft - 1 - ich;
    const int sh = dedisp_shift(row, nrows, msamp, dm, fmin, foff, tsamp_wf);
    const int nbin = N * len_sft;
    const int64_t seg_off = (int64_t)c * nbin + (int64_t)k * N + t;
    float2 d[8];
    if (fac) {
#pragma unroll
        for (int m = 0; m < 8; ++m) d[m] = fac[seg_off + T * m];
    } else {
        // segment k holds sub-band ich's bins (the half-row rotation)
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            int col;
            d[m] = chirp_factor(c, ich, t + T * m, len_sft, N, dm, fmin, ca.bw_sub, ca.bw_chan, ca.taper, ca.inv_nbin,
                                &col);
        }
    }
    const int64_t plane = (int64_t)nfft * nchan * nbin;
    const int b0 = (int)blockIdx.y * bpw;
    const int b1 = min(nfft, b0 + bpw);
    float* out = wf + (int64_t)row * msamp;
    __syncthreads();  // the twiddle table
    // iterations (block b, polarisation p) in order, the next one's spectra
    // loaded while this one's FFT runs
    const int nit = (b1 - b0) * npairs;
    const int64_t bstride = (int64_t)nchan * nbin;
    const float2* src = spec + (int64_t)b0 * bstride + seg_off;
    float2 zn[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) zn[m] = src[T * m];
    float v[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) v[m] = 0.0f;
    int b = b0, p = 0;
    for (int it = 0; it < nit; ++it) {
        float2 z[8];
#pragma unroll
        for (int m = 0; m < 8; ++m) z[m] = zn[m];
        if (it + 1 < nit) {
            const float2* nsrc = p + 1 < npairs ? src + (int64_t)(p + 1) * plane : src + bstride;
#pragma unroll
            for (int m = 0; m < 8; ++m) zn[m] = nsrc[T * m];
        }
#pragma unroll
        for (int m = 0; m < 8; ++m) z[m] = cmul_fma(z[m], d[m]);
        fft_passes<N, 1, 0>(z, lds, tw, t);
        // |x|^2 / nbin^2 added over the polarisations, in float
#pragma unroll
        for (int m = 0; m < 8; ++m) v[m] = fmaf(fmaf(z[m].x, z[m].x, z[m].y * z[m].y), scale, v[m]);
        if (++p == npairs) {
#pragma unroll
            for (int m = 0; m < 8; ++m) {
                const int n = t + T * m;
                if (n >= nov_sub && n < N - nov_sub) {
                    const int s0 = b * mbin_adj + (n - nov_sub);  // + sh, modulo msamp
                    out[s0 >= msamp - sh ? s0 - (msamp - sh) : s0 + sh] = v[m];
                }
                v[m] = 0.0f;
            }
            p = 0;
            ++b;
            src += bstride;
        }
    }
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
    DevBuf chirped;  // the chirped spectra, then their sub-band time series (unfused path)
    DevBuf factors;  // one DM's chirp factors [chan][nbin] (fused path reading a table)
    DevBuf taper;    // the sub-band taper per bin, mbin doubles
    DevBuf twiddles; // the fused pass's FFT twiddle table (fft_twiddle_table)
    double bw_sub = 0, bw_chan = 0, inv_nbin = 0;
    ~sps_coherent() {
        if (have_fwd) (void)hipfftDestroy(fwd);
        if (have_inv) (void)hipfftDestroy(inv);
    }
};

namespace {

void coherent_chirp(const sps_coherent& h, const float* d_in, double dm, float* d_out, hipStream_t st) {
    const CoherentShape& s = h.sh;
    const dim3 grid((unsigned)ceil_div((int64_t)s.mbin(), (int64_t)kThreads), (unsigned)s.len_sft,
                    (unsigned)s.nchan);
    hipLaunchKernelGGL(k_coherent_chirp, grid, dim3(kThreads), 0, st, reinterpret_cast<const float2*>(d_in),
                       reinterpret_cast<float2*>(d_out), dm, s.nchan, s.len_sft, s.mbin(), s.nfft * (s.npol / 2),
                       s.fmin, h.bw_sub, h.bw_chan, h.taper.as<double>(), h.inv_nbin);
    SPS_HIP(hipGetLastError());
}

// The twiddle table of the fused pass's N-point inverse FFT, in the layout
// FftPlan<N> reads: for each pass after the first (stride ns, radix R), the
// (R - 1) * ns factors exp(2 pi i r kk / (ns R)), r-major.
std::vector<float2> fft_twiddle_table(int n) {
    std::vector<float2> tw;
    const double pi = 3.14159265358979323846;
    for (int ns = 8; ns < n;) {
        const int r_ = (n / ns) >= 8 ? 8 : n / ns;
        for (int e = 0; e < (r_ - 1) * ns; ++e) {
            const int r = e / ns + 1, kk = e % ns;
            const double a = 2.0 * pi * (double)(r * kk) / (double)(ns * r_);
            tw.push_back(make_float2((float)std::cos(a), (float)std::sin(a)));
        }
        ns *= r_;
    }
    return tw;
}

// the fused pass applies: a power-of-two sub-band length the kernel is built
// for, and an even sub-band count (the half-row rotation moves whole
// segments)
bool coherent_fusable(const CoherentShape& s) {
    const int n = s.mbin();
    return s.len_sft % 2 == 0 && n >= 128 && n <= 4096 && (n & (n - 1)) == 0;
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

// the fused pass computes each segment's chirp factors itself rather than
// reading a per-DM factor table
constexpr bool kFusedOwnFactors = true;

template <int N>
void launch_fused(const CoherentShape& s, const float2* spec, const float2* fac, const ChirpArgs& ca,
                  const float2* twg, float* wf, double dm, int bpw, double foff, double tsamp_wf, float scale,
                  hipStream_t st) {
    static_assert(FftPlan<N>::tw_size > 0, "at least two passes");
    const dim3 grid((unsigned)s.rows(), (unsigned)ceil_div((int64_t)s.nfft, (int64_t)bpw));
    hipLaunchKernelGGL(k_coherent_fused<N>, grid, dim3(N / 8), 0, st, spec, fac, ca, twg, wf, s.npol / 2, s.nchan,
                       s.len_sft,
                       s.nfft, bpw, s.nov_sub(), s.mbin_adj(), (int)s.msamp(), dm, s.fmin, foff, tsamp_wf, scale);
}

void coherent_fused(sps_coherent& h, const float* d_spectra, double dm, float* d_wf, hipStream_t st);

void coherent_fft(sps_coherent& h, hipfftHandle plan, float* d, int dir, hipStream_t st) {
    check_hipfft(hipfftSetStream(plan, st), "set stream");
    check_hipfft(hipfftExecC2C(plan, reinterpret_cast<hipfftComplex*>(d), reinterpret_cast<hipfftComplex*>(d), dir),
                 "C2C");
}

void coherent_fused(sps_coherent& h, const float* d_spectra, double dm, float* d_wf, hipStream_t st) {
    const CoherentShape& s = h.sh;
    const ChirpArgs ca{h.bw_sub, h.bw_chan, h.taper.as<double>(), h.inv_nbin};
    float2* fac = nullptr;  // the fused pass computes its factors
    if (!kFusedOwnFactors) {
        fac = h.factors.as<float2>();
        const dim3 fgrid((unsigned)ceil_div((int64_t)s.mbin(), (int64_t)kThreads), (unsigned)s.len_sft,
                         (unsigned)s.nchan);
        hipLaunchKernelGGL(k_coherent_factors, fgrid, dim3(kThreads), 0, st, fac, dm, s.len_sft, s.mbin(), s.fmin,
                           h.bw_sub, h.bw_chan, h.taper.as<double>(), h.inv_nbin);
        SPS_HIP(hipGetLastError());
    }
    const double foff = (s.fmax - s.fmin) / ((double)s.nchan * s.len_sft);
    const double tsamp_wf = (double)((float)(unsigned)s.len_sft * s.tsamp);
    const float scale = (float)(1.0 / ((double)s.nbin * (double)s.nbin));
    // blocks per launch block: enough launch blocks to fill the chip several
    // times over, each reusing its segment's factors for the rest
    const int64_t want = 2048;
    const int bpw = (int)std::max<int64_t>(1, std::min<int64_t>(s.nfft, ceil_div(s.rows() * s.nfft, want)));
    const float2* spec = reinterpret_cast<const float2*>(d_spectra);
    const float2* twg = h.twiddles.as<float2>();
    switch (s.mbin()) {
        case 128: launch_fused<128>(s, spec, fac, ca, twg, d_wf, dm, bpw, foff, tsamp_wf, scale, st); break;
        case 256: launch_fused<256>(s, spec, fac, ca, twg, d_wf, dm, bpw, foff, tsamp_wf, scale, st); break;
        case 512: launch_fused<512>(s, spec, fac, ca, twg, d_wf, dm, bpw, foff, tsamp_wf, scale, st); break;
        case 1024: launch_fused<1024>(s, spec, fac, ca, twg, d_wf, dm, bpw, foff, tsamp_wf, scale, st); break;
        case 2048: launch_fused<2048>(s, spec, fac, ca, twg, d_wf, dm, bpw, foff, tsamp_wf, scale, st); break;
        case 4096: launch_fused<4096>(s, spec, fac, ca, twg, d_wf, dm, bpw, foff, tsamp_wf, scale, st); break;
        default: throw Error(SPS_ERR_RUNTIME, "fused coherent pass: unsupported sub-band length");
    }
    SPS_HIP(hipGetLastError());
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
        if (coherent_fusable(h->sh) && !kFusedOwnFactors)
            h->factors.ensure(sizeof(float) * 2 * (size_t)nchan * nbin);
        // double bw = m_Fmax - m_Fmin: a float difference (Chunk_fly_gpu.cu:96)
        const double bw = (double)(f_max - f_min);
        h->bw_sub = bw / nchan;
        h->bw_chan = h->bw_sub / len_sft;
        h->inv_nbin = (nbin & (nbin - 1)) == 0 ? 1.0 / nbin : 0.0;
        const int mbin = h->sh.mbin();
        std::vector<double> taper((size_t)mbin);
        for (int i = 0; i < mbin; ++i) taper[(size_t)i] = chirp_taper(i, mbin, h->bw_chan);
        h->taper.ensure(sizeof(double) * (size_t)mbin);
        SPS_HIP(hipMemcpy(h->taper.p, taper.data(), sizeof(double) * (size_t)mbin, hipMemcpyHostToDevice));
        if (coherent_fusable(h->sh)) {
            const std::vector<float2> tw = fft_twiddle_table(mbin);
            h->twiddles.ensure(sizeof(float2) * tw.size());
            SPS_HIP(hipMemcpy(h->twiddles.p, tw.data(), sizeof(float2) * tw.size(), hipMemcpyHostToDevice));
        }
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

SPS_API sps_status sps_coherent_is_fused(const sps_coherent* h, int* h_fused) {
    return guarded([&] {
        SPS_REQUIRE(h && h_fused, "null argument");
        *h_fused = coherent_fusable(h->sh) ? 1 : 0;
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
        if (coherent_fusable(h->sh)) {
            coherent_fused(*h, d_spectra, dm, d_waterfall, st);
            return;
        }
        h->chirped.ensure(sizeof(float) * 2 * (size_t)h->sh.spectra());
        float* buf = h->chirped.as<float>();
        coherent_chirp(*h, d_spectra, dm, buf, st);
        coherent_fft(*h, h->inv, buf, HIPFFT_BACKWARD, st);
        coherent_detect(*h, buf, dm, d_waterfall, st);
    });
}

}  // extern "C"
