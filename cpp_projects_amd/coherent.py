"""Python mirror of the coherent front-end over the C ABI: CChunk_gpu::process
steps 2-8 with CChunk_fly_gpu's on-the-fly chirp (CDMT/Chunk_gpu.cu:197-209,
254-322; Chunk_fly_gpu.cu:94-164).  Spectra are complex64 device tensors laid
out [pol][fft][chan][bin] (the unpack's output); the waterfall is a float32
device tensor [nchan * len_sft][msamp] that the search takes as its input.

CoherentChunk.process runs the whole of CChunk_gpu::process on one chunk of
unpacked blocks: the forward FFT, then per coherent DM the waterfall, the
statistics and normalisation, the CFdmt of the keep mask and of the waterfall
and the S/N normalisation into that DM's rows of the output image, and at the
end the clustering over the image with the candidates log.
"""
import ctypes as C

from . import _lib
from ._lib import call
from .fdmt import _dptr, _stream


class CoherentChunk:
    """The front-end of one chunk geometry (CChunk_gpu's constructor
    arguments that shape it: band, polarisation count, coarse channels,
    sub-channels per channel, FFT block length, blocks, overlap and the raw
    sampling time)."""

    def __init__(self, f_min, f_max, npol, nchan, len_sft, nbin, nfft, noverlap, tsamp):
        self.f_min, self.f_max, self.len_sft, self.tsamp = float(f_min), float(f_max), int(len_sft), float(tsamp)
        self._search = None
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
        if h is not None and h.value:
            try:
                _lib.lib().sps_coherent_destroy(h)
            except Exception:  # interpreter shutdown: the library may be gone
                pass
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

    def process(self, d_blocks, dms, sigma_bound, d_image=None, wnd=4, bins=(2, 1, 2), filename="candidates.log",
                t_chunk_begin=0.0, bwrite=True):
        """CChunk_gpu::process (CDMT/Chunk_gpu.cu:254-425) on unpacked blocks,
        which the forward FFT overwrites with their spectra (step 2).  For
        each coherent DM idm (steps 3-14): the waterfall, then one search call
        over the CFdmt plan (nchan * len_sft channels, msamp columns, imaxDt =
        len_sft; sps_search_create_cfdmt) that does calc_fdmt_inputs' statistics,
        the transform of the keep mask (over its causal extent: the mask is
        constant along each row), the transform of the waterfall with the
        normalisation inside the init's loads, and fdmt_normalization inside
        the last pass, into rows [idm * len_sft, (idm + 1) * len_sft) of the
        image -- the same values as the reference's separate steps.  Then clusterization_main over the image (steps 16-17:
        widths 1..wnd, threshold sigma_bound, gap bins, the log written to
        filename with the waterfall's sampling time len_sft * tsamp).  The
        reference's power-of-two gather before the clustering keeps no
        result and is not run.

        Returns (image, candidates, delegates, group sizes): the float32
        image [len(dms) * len_sft][msamp], the all-widths candidates and
        the delegate indices into them with their group sizes (empty when
        no value passes the threshold)."""
        import torch

        from .search import SingleBeamSearch, _write_delegates, clusterization

        dev = d_blocks.device
        rows, msamp, L = self.rows, self.cols, self.len_sft
        ndm = len(dms)
        if ndm == 0:
            raise ValueError("process needs at least one coherent DM")
        if d_image is None:
            d_image = torch.empty((ndm * L, msamp), dtype=torch.float32, device=dev)
        self._check(d_image, ndm * L * msamp, "image")
        if self._search is None:
            # its own gather (widths 1, nothing stored) is not used
            self._search = SingleBeamSearch(self.f_min, self.f_max, rows, msamp, L - 1, in_dtype="f32", box_q=0,
                                            threshold=1e30, cand_capacity=0, cfdmt_imaxdt=L)
            self._wf = torch.empty((rows, msamp), dtype=torch.float32, device=dev)
            self._cands = torch.empty((1, 4), dtype=torch.int32, device=dev)
            self._count = torch.zeros(1, dtype=torch.int64, device=dev)
        img = d_image.view(ndm * L, msamp)
        self.forward(d_blocks)
        for idm, dm in enumerate(dms):
            self.waterfall(d_blocks, dm, self._wf)
            self._search.run(self._wf, img[idm * L:(idm + 1) * L], self._cands, self._count)
        d_thr = torch.tensor([float(sigma_bound)], dtype=torch.float32, device=dev)
        d_bins = torch.tensor(list(bins), dtype=torch.int32, device=dev)
        ok, cands, n = clusterization.gather_candidates_in_heap_with_plan(img, ndm * L, msamp, d_thr, wnd)
        if not ok:
            empty = torch.empty(0, dtype=torch.int64)
            return d_image, cands, empty, empty
        dele, gsz = clusterization.cluster(cands, n, d_bins)
        if bwrite:
            _write_delegates(cands, n, dele.cpu().numpy(), gsz.cpu().numpy(), filename, L * self.tsamp, t_chunk_begin)
        return d_image, cands, dele, gsz
