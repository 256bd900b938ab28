"""Batched multi-beam search over the C ABI (sps_batch_*).

Many independent beams, each with its own channel count, length and maximum
delay, run the single-beam search (normalise -> FDMT -> S/N normalisation ->
boxcar gather) in one launch sequence.  The library spreads the beams over a
few streams, shares scratch states between the beams of one stream, and
captures the sequence into a HIP graph that later calls replay.  The device
tensors handed in are bound to the batch; the object keeps them alive.
"""
import ctypes as C

from . import _lib
from ._lib import SPS_NORM_GLOBAL, BatchBeam, BatchParams, call
from .fdmt import _dptr, _stream


class Beam:
    """One beam of a batch.  d_dmt may be None (the image then stays in
    scratch); d_cands is an int32 (capacity, 4) tensor; d_count one int64."""

    def __init__(self, f_min, f_max, nchans, nsamps, dt_max, d_in, d_cands, d_count, d_dmt=None, tsamp=1.0):
        self.f_min, self.f_max, self.tsamp = float(f_min), float(f_max), float(tsamp)
        self.nchans, self.nsamps, self.dt_max = int(nchans), int(nsamps), int(dt_max)
        self.d_in, self.d_dmt, self.d_cands, self.d_count = d_in, d_dmt, d_cands, d_count
        if d_in.numel() != self.nchans * self.nsamps:
            raise ValueError("input size does not match nchans x nsamps")
        self.cand_capacity = int(d_cands.shape[0]) if d_cands is not None else 0


class BatchSearch:
    def __init__(self, beams, in_dtype="f32", norm_mode=SPS_NORM_GLOBAL, snr_norm=True, box_q=8, threshold=6.0,
                 nworkers=4):
        self.beams = list(beams)
        p = BatchParams()
        p.in_dtype = _lib.SPS_U8 if in_dtype == "u8" else _lib.SPS_F32
        p.norm_mode, p.snr_norm, p.box_q = int(norm_mode), int(snr_norm), int(box_q)
        p.threshold, p.nworkers = float(threshold), int(nworkers)
        self.params = p
        arr = (BatchBeam * len(self.beams))()
        for i, b in enumerate(self.beams):
            e = arr[i]
            e.f_min, e.f_max, e.tsamp = b.f_min, b.f_max, b.tsamp
            e.nchans, e.nsamps, e.dt_max = b.nchans, b.nsamps, b.dt_max
            e.d_in = _dptr(b.d_in)
            e.d_dmt = _dptr(b.d_dmt) if b.d_dmt is not None else None
            e.d_cands = _dptr(b.d_cands) if b.d_cands is not None else None
            e.cand_capacity = b.cand_capacity
            e.d_count = _dptr(b.d_count)
        self._h = C.c_void_p()
        call("sps_batch_create", C.byref(p), arr, len(self.beams), C.byref(self._h))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.lib().sps_batch_destroy(h)
            except Exception:
                pass
            self._h = None

    def beam_info(self, i):
        """(ndt, worker stream index, FDMT iterations) of beam i."""
        ndt, w, n = C.c_int64(), C.c_int32(), C.c_int64()
        call("sps_batch_get_beam_info", self._h, int(i), C.byref(ndt), C.byref(w), C.byref(n))
        return ndt.value, w.value, n.value

    def set_schedule(self, i, depths):
        arr = (C.c_int32 * max(len(depths), 1))(*depths)
        call("sps_batch_set_schedule", self._h, int(i), arr, len(depths))

    def set_staged(self, i, mode=1):
        call("sps_batch_set_staged", self._h, int(i), int(mode))

    def set_group_shape(self, i, lds_floats=8192, max_rows=16, tiles_per_block=8):
        call("sps_batch_set_group_shape", self._h, int(i), int(lds_floats), int(max_rows), int(tiles_per_block))

    def get_schedule(self, i):
        n = C.c_int64()
        buf = (C.c_int32 * 64)()
        call("sps_batch_get_schedule", self._h, int(i), buf, 64, C.byref(n))
        return list(buf)[: n.value]

    def run(self):
        """Enqueues every beam directly on the current stream's fork/join."""
        call("sps_batch_run", self._h, _stream())

    def launch(self):
        """Replays the captured graph (captures it on first use)."""
        call("sps_batch_launch", self._h, _stream())

    def sort_candidates(self, i):
        n = C.c_int64()
        call("sps_batch_sort_candidates", self._h, int(i), C.byref(n), _stream())
        return n.value
