"""Python mirror of the coherent front-end over the C ABI: CChunk_gpu::process
steps 2-8 with CChunk_fly_gpu's on-the-fly chirp (CDMT/Chunk_gpu.cu:197-209,
254-322; Chunk_fly_gpu.cu:94-164).  Spectra are complex64 device tensors laid
out [pol][fft][chan][bin] (the unpack's output); the waterfall is a float32
device tensor [nchan * len_sft][msamp] that the search takes as its input.
"""
import ctypes as C

from ._lib import call
from .fdmt import _dptr, _stream


class CoherentChunk:
    """The front-end of one chunk geometry (CChunk_gpu's constructor
    arguments that shape it: band, polarisation count, coarse channels,
    sub-channels per channel, FFT block length, blocks, overlap and the raw
    sampling time)."""

    def __init__(self, f_min, f_max, npol, nchan, len_sft, nbin, nfft, noverlap, tsamp):
        h = C.c_void_p()
        call("sps_coherent_create", float(f_min), float(f_max), int(npol), int(nchan), int(len_sft), int(nbin),
             int(nfft), int(noverlap), float(tsamp), C.byref(h))
        self._h = h
        r, c, n = C.c_int64(), C.c_int64(), C.c_int64()
        call("sps_coherent_get_shape", h, C.byref(r), C.byref(c), C.byref(n))
        self.rows, self.cols, self.spectra = r.value, c.value, n.value
        f = C.c_int()
        call("sps_coherent_is_fused", h, C.byref(f))
        # waterfall() runs one fused pass per DM (power-of-two sub-band length
        # 128..4096, even sub-band count) rather than chirp, inverse FFT and
        # intensity pass
        self.fused = bool(f.value)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            call("sps_coherent_destroy", h)
            self._h = None

    def _check(self, t, n, what):
        if t.numel() * (2 if t.is_complex() else 1) < n:
            raise ValueError(f"{what} holds fewer than {n} floats")

    def forward(self, d_blocks):
        """The forward FFT of every block, in place (Chunk_gpu.cu:267)."""
        self._check(d_blocks, 2 * self.spectra, "blocks")
        call("sps_coherent_forward", self._h, _dptr(d_blocks), _stream())

    def waterfall(self, d_spectra, dm, d_waterfall):
        """Chirp, inverse FFT and intensity waterfall for one coherent DM
        (Chunk_gpu.cu:276-322), fused into one pass where `fused`."""
        self._check(d_spectra, 2 * self.spectra, "spectra")
        self._check(d_waterfall, self.rows * self.cols, "waterfall")
        call("sps_coherent_waterfall", self._h, _dptr(d_spectra), float(dm), _dptr(d_waterfall), _stream())

    def chirp(self, d_spectra, dm, d_out):
        """kernel_el_wise_mult_onthe_fly (Chunk_fly_gpu.cu:119-164)."""
        self._check(d_spectra, 2 * self.spectra, "spectra")
        self._check(d_out, 2 * self.spectra, "output")
        call("sps_coherent_chirp", self._h, _dptr(d_spectra), float(dm), _dptr(d_out), _stream())

    def inverse(self, d_buf):
        """The inverse FFT of every sub-band, in place (Chunk_gpu.cu:288)."""
        self._check(d_buf, 2 * self.spectra, "buffer")
        call("sps_coherent_inverse", self._h, _dptr(d_buf), _stream())

    def detect(self, d_buf, dm, d_waterfall):
        """roll_rows_normalize_sum_kernel, transpose_unpadd_intensity and
        dedisperse (Chunk_gpu.cu:476-515,613-632) in one pass."""
        self._check(d_buf, 2 * self.spectra, "buffer")
        self._check(d_waterfall, self.rows * self.cols, "waterfall")
        call("sps_coherent_detect", self._h, _dptr(d_buf), float(dm), _dptr(d_waterfall), _stream())
