"""Python mirror of the reference normalisation, detection and clustering
calls over the C ABI (statistical_preprocessing::, fdmt_normalization,
clusterization::, select_delegates::, doDetectSignal).  Device buffers are
torch tensors; candidate arrays are int32 tensors of shape (n, 4) holding
Cand{mt, mdt, mwidth, msnr} (Candidate.cuh:4-10) with msnr's bits in
column 3.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import SPS_NORM_GLOBAL, SPS_NORM_NONE, SPS_NORM_PER_CHANNEL, SearchParams, call  # noqa: F401
from .fdmt import _dptr, _stream

CAND_DTYPE = np.dtype([("mt", np.int32), ("mdt", np.int32), ("mwidth", np.int32), ("msnr", np.float32)])


def cands_to_numpy(t, n=None):
    """(n, 4) int32 device/host tensor -> numpy structured array."""
    a = t.detach().cpu().numpy() if hasattr(t, "detach") else np.asarray(t)
    if n is not None:
        a = a[:n]
    return np.ascontiguousarray(a, np.int32).view(CAND_DTYPE).reshape(-1)


def numpy_to_cands(c, device="cuda"):
    import torch

    a = np.ascontiguousarray(c, CAND_DTYPE).view(np.int32).reshape(-1, 4)
    return torch.from_numpy(a.copy()).to(device)


class statistical_preprocessing:
    @staticmethod
    def calc_fdmt_inputs(d_in, nrows, ncols, d_buff, d_norm_input, mode=SPS_NORM_GLOBAL):
        """statistical_preprocessing::calc_fdmt_inputs (Statistical_preprocessing.cuh:9-10)."""
        call("sps_calc_fdmt_inputs", _dptr(d_in), int(nrows), int(ncols), _dptr(d_buff), _dptr(d_norm_input),
             int(mode), _stream())

    @staticmethod
    def clean_fdmtInput_and_create_normInput(d_arr, nrows, ncols, d_pbarr_norm, d_buff):
        """statistical_preprocessing::clean_fdmtInput_and_create_normInput
        (Statistical_preprocessing.cuh:6-7): rows whose variance is more than
        4 sigma from the mean row variance are zeroed in place; d_pbarr_norm
        (a bool tensor of nrows*ncols, or None) receives the keep mask;
        d_buff (>= nrows + 2 floats, or None) the row variances, their mean
        and standard deviation."""
        call("sps_clean_fdmt_input", _dptr(d_arr), int(nrows), int(ncols), _dptr(d_pbarr_norm), _dptr(d_buff),
             _stream())

    @staticmethod
    def calcThreadsForMean_and_Disp(n_cols):
        """Statistical_preprocessing.cuh:12-17: the power of two >= n_cols, at most 1024."""
        k = 0
        while (1 << k) < n_cols and k < 10:
            k += 1
        return 1 << k


def fnc_cumsum_customized_gpu(d_in, d_buff, length, d_out):
    """Inclusive int32 prefix sums (cumsum_customized.cu:158-179); d_buff is
    not used (the scratch is the library's)."""
    call("sps_scan_inclusive_i32", _dptr(d_in), _dptr(d_out), int(length), _stream())


def inclusiveScan_CUB(d_arr, length):
    """In-place inclusive prefix sums of an int32 or float32 tensor
    (cumsum_customized.cu:259-305)."""
    name = "sps_scan_inclusive_f32" if d_arr.dtype.is_floating_point else "sps_scan_inclusive_i32"
    call(name, _dptr(d_arr), _dptr(d_arr), int(length), _stream())


def fdmt_normalization(d_arr, d_norm, length, d_out):
    """fdmt_normalization (CDMT/Chunk_gpu.cu:841-852)."""
    call("sps_fdmt_normalization", _dptr(d_arr), _dptr(d_norm), int(length), _dptr(d_out), _stream())


def fdmt_normalization_rows(d_arr, rows, cols, d_norm, norm_cols, d_out):
    """fdmt_normalization over a rows x cols image with a norm image given
    for its first norm_cols columns and constant along each row beyond:
    d_out[r][t] = d_arr[r][t] / sqrtf(d_norm[r][min(t, norm_cols - 1)] + 1e-8)."""
    call("sps_fdmt_normalization_rows", _dptr(d_arr), int(rows), int(cols), _dptr(d_norm), int(norm_cols),
         _dptr(d_out), _stream())


def _gather(d_fdmt, rows, cols, d_thr, wnd, all_widths, capacity=None):
    import torch

    cap = int(capacity) if capacity is not None else 1 << 16
    while True:
        out = torch.empty((max(cap, 1), 4), dtype=torch.int32, device=d_fdmt.device)
        n = C.c_int64()
        st = _lib.lib().sps_boxcar_candidates(_dptr(d_fdmt), int(rows), int(cols), _dptr(d_thr), int(wnd),
                                              int(all_widths), _dptr(out), cap, C.byref(n), _stream())
        if st == _lib.SPS_ERR_CAPACITY and capacity is None:
            cap = n.value
            continue
        _lib.check(st)
        return out[: n.value], n.value


class clusterization:
    @staticmethod
    def gather_candidates_in_dynamicalArray(d_arr_fdmt, rows, cols, val_thresh, quant_power2_wnd):
        """Clusterization.cuh:22-23: widths 1..2^q, mwidth = log2(w).  Returns
        (candidates (n, 4) int32 device tensor, n); order (mdt, mt, mwidth)."""
        return _gather(d_arr_fdmt, rows, cols, val_thresh, quant_power2_wnd, 0)

    @staticmethod
    def gather_candidates_in_heap_with_plan(d_fdmt, rows, cols, d_val_thresh, wnd_width):
        """Clusterization.cuh:56-63: widths 1..W, mwidth = w, row-major order.
        Returns (ok, candidates, n) with ok False when there are none."""
        c, n = _gather(d_fdmt, rows, cols, d_val_thresh, wnd_width, 1)
        return n > 0, c, n

    @staticmethod
    def digitize_kernel(d_arrfdmt, len_arr=None):
        """Clusterization.cu:206-214: d_arrfdmt[i] = floorf(d_arrfdmt[i]) in
        place for i < len_arr (default: the whole contiguous f32 tensor)."""
        n = d_arrfdmt.numel() if len_arr is None else int(len_arr)
        if n > d_arrfdmt.numel():
            raise ValueError("len_arr exceeds the tensor")
        call("sps_digitize", _dptr(d_arrfdmt), n, _stream())

    @staticmethod
    def cluster(d_cands, n, d_bins):
        """Three-key gap clustering + delegates (Clusterization.cu:92-166).
        Returns (delegates int32 device tensor sorted by mt, group sizes int64)."""
        import torch

        dev = d_cands.device
        dele = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
        gsz = torch.empty(max(n, 1), dtype=torch.int64, device=dev)
        ng = C.c_int64()
        call("sps_cluster", _dptr(d_cands) if n else None, int(n), _dptr(d_bins), _dptr(dele), _dptr(gsz),
             C.byref(ng), _stream())
        return dele[: ng.value], gsz[: ng.value]

    @staticmethod
    def complete_clusterization(d_cands, n, d_bins, filename, delt_bin, t_chunk_begin, bwrite=True):
        """Clusterization.cuh:25-32: gap clustering by (mt, mdt, mwidth) with
        d_bins, delegate selection, delegates sorted by mt and the log.  Like
        the reference it returns nothing; cluster() returns the delegates."""
        dele, gsz = clusterization.cluster(d_cands, n, d_bins)
        if bwrite:
            _write_delegates(d_cands, n, dele.cpu().numpy(), gsz.cpu().numpy(), filename, delt_bin, t_chunk_begin)

    @staticmethod
    def clusterization_main(d_fdmt, rows, cols, d_val_thresh, wnd_width, d_bins, filename, delt_bin, t_chunk_begin,
                            bwrite=True):
        """Clusterization.cuh:10-20: all-widths gather, clustering, log.
        Returns False when no candidate passes the threshold."""
        ok, c, n = clusterization.gather_candidates_in_heap_with_plan(d_fdmt, rows, cols, d_val_thresh, wnd_width)
        if not ok:
            return False
        clusterization.complete_clusterization(c, n, d_bins, filename, delt_bin, t_chunk_begin, bwrite)
        return True

    @staticmethod
    def writeCandHostArrayToCSV(h_cands, filename):
        """Clusterization.cu:338-358: header mt,mdt,mwidth,msnr."""
        with open(filename, "w") as f:
            f.write("mt,mdt,mwidth,msnr\n")
            for c in h_cands:
                f.write(f"{int(c['mt'])},{int(c['mdt'])},{int(c['mwidth'])},{float(c['msnr']):g}\n")


def _write_delegates(d_cands, n, delegates, group_sizes, filename, delt_bin, t_chunk_begin):
    dl = np.ascontiguousarray(delegates, np.int32)
    gs = np.ascontiguousarray(group_sizes, np.int64)
    call("sps_write_log", _dptr(d_cands) if n else None, int(n), dl.ctypes.data_as(C.c_void_p),
         gs.ctypes.data_as(C.c_void_p), len(dl), str(filename).encode(), float(delt_bin), float(t_chunk_begin))


class select_delegates:
    @staticmethod
    def select_delegates(d_cands, d_indices, d_group_begin, d_out=None):
        """select_delegates::select_delegates (select_delegates.cu:16-51):
        for group g = positions [begin[g], begin[g+1]) of d_indices (the last
        ends at len(d_indices)), the candidate index with the largest msnr,
        the first position on ties.  Returns an int32 device tensor."""
        import torch

        ng = int(d_group_begin.numel())
        out = d_out if d_out is not None else torch.empty(max(ng, 1), dtype=torch.int32, device=d_cands.device)
        call("sps_select_delegates", _dptr(d_cands), _dptr(d_indices), int(d_indices.numel()), _dptr(d_group_begin),
             ng, _dptr(out), _stream())
        return out[:ng]

    @staticmethod
    def select_delegates_kernel(d_cands, d_indices, quant_cand, d_group_begin, d_num_delegates, quant_groups):
        """select_delegates_kernel (select_delegates.cu:62-125), launched by the
        reference with one block per group: quant_cand ends the last group."""
        call("sps_select_delegates", _dptr(d_cands), _dptr(d_indices), int(quant_cand), _dptr(d_group_begin),
             int(quant_groups), _dptr(d_num_delegates), _stream())

    @staticmethod
    def write_log(d_cands, quant_candidates, d_group_begin, d_indices, d_num_delegates, filename, delt_bin,
                  t_chunk_begin):
        """select_delegates::write_log (select_delegates.cu:127-177): delegate
        i is d_num_delegates[i] and its num column is the size of group i."""
        begin = d_group_begin.cpu().numpy().astype(np.int64)
        ends = np.append(begin[1:], int(d_indices.numel()))
        _write_delegates(d_cands, quant_candidates, d_num_delegates.cpu().numpy(), ends - begin, filename,
                         delt_bin, t_chunk_begin)


def doDetectSignal_v1(d_arr, d_norm, rows, cols, wnd, thr, grid, block, sz, d_aux, d_aux_num, d_width, out,
                      version=1):
    """doDetectSignal_v1 (detection/kernel.cu:561-582): block[0] columns per
    tile, grid (ceil(cols/block[0]), rows); d_aux (float), d_aux_num and
    d_width (int32) hold one entry per tile; out a (4,) int32 device tensor
    receiving {iwidth, irow, icol, snr bits}.  Returns snr >= thr."""
    tile = int(block[0])
    if int(grid[0]) != (cols + tile - 1) // tile or int(grid[1]) != rows:
        raise _lib.InvalidArgument(_lib.SPS_ERR_INVALID_ARGUMENT, "grid must be (ceil(cols/block.x), rows)")
    call("sps_detect_signal", int(version), _dptr(d_arr), _dptr(d_norm), int(rows), int(cols), int(wnd), tile,
         _dptr(d_aux), _dptr(d_aux_num), _dptr(d_width), int(d_aux.numel()), _dptr(out), _stream())
    o = out.cpu()
    import torch

    return float(o[3:4].view(torch.float32)[0]) >= thr


def doDetectSignal_v2(d_arr, d_norm, rows, cols, wnd, thr, grid, block, sz, d_aux, d_aux_num, d_width, out):
    """doDetectSignal_v2 (detection/kernel.cu:584-605): energy ratio
    sqrt(sum x^2 / sum n^2) over windows of 1..wnd samples."""
    return doDetectSignal_v1(d_arr, d_norm, rows, cols, wnd, thr, grid, block, sz, d_aux, d_aux_num, d_width, out,
                             version=2)


def doDetectSignal_v2_2(d_arr, d_norm, rows, cols, wnd, thr, grid, block, sz, d_aux, d_aux_num, d_width, out,
                        arr_buff=None):
    """doDetectSignal_v2_2 (detection/kernel.cu:609-640): the energy ratio of
    doDetectSignal_v2 with a one-block final reduction (its window kernel is
    declared in a header the reference does not ship)."""
    return doDetectSignal_v1(d_arr, d_norm, rows, cols, wnd, thr, grid, block, sz, d_aux, d_aux_num, d_width, out,
                             version=2)


def detect_max(d_arr, d_norm, rows, cols, wnd):
    """doDetectSignal_v1 (detection/kernel.cu:561-582) -> (iwidth, irow, icol, snr)."""
    import torch

    out = torch.zeros(4, dtype=torch.int32, device=d_arr.device)
    call("sps_detect_max", _dptr(d_arr), _dptr(d_norm), int(rows), int(cols), int(wnd), _dptr(out), _stream())
    o = out.cpu()
    return int(o[0]), int(o[1]), int(o[2]), float(o[3:4].view(torch.float32)[0])


def boxcar_snr(d_img, rows, cols, w, d_snr):
    call("sps_boxcar_snr", _dptr(d_img), int(rows), int(cols), int(w), _dptr(d_snr), _stream())


def boxcar_snr_adjoint(d_g, rows, cols, w, d_gi):
    call("sps_boxcar_snr_adjoint", _dptr(d_g), int(rows), int(cols), int(w), _dptr(d_gi), _stream())


class SingleBeamSearch:
    """The CChunk_gpu::process hot section (CDMT/Chunk_gpu.cu:325-425) for one
    beam: normalise -> FDMT -> S/N normalisation -> boxcar gather."""

    def __init__(self, f_min, f_max, nchans, nsamps, dt_max, tsamp=1.0, in_dtype="u8", norm_mode=SPS_NORM_GLOBAL,
                 snr_norm=True, box_q=8, threshold=6.0, cand_capacity=1 << 20, cfdmt_imaxdt=None):
        """cfdmt_imaxdt: run the legacy CFdmt plan (CFdmtGpu(f_min, f_max,
        nchans, nsamps, cfdmt_imaxdt)) instead of the FDMT plan, as
        CChunk_gpu::process does per coherent DM; dt_max and tsamp are then
        unused."""
        p = SearchParams()
        p.f_min, p.f_max, p.tsamp = f_min, f_max, tsamp
        p.nchans, p.nsamps, p.dt_max = nchans, nsamps, dt_max
        p.in_dtype = _lib.SPS_U8 if in_dtype == "u8" else _lib.SPS_F32
        p.norm_mode, p.snr_norm, p.box_q = norm_mode, int(snr_norm), box_q
        p.threshold, p.cand_capacity = threshold, cand_capacity
        self.params = p
        self._h = C.c_void_p()
        if cfdmt_imaxdt is None:
            call("sps_search_create", C.byref(p), C.byref(self._h))
        else:
            call("sps_search_create_cfdmt", C.byref(p), int(cfdmt_imaxdt), C.byref(self._h))
        n = C.c_int64()
        call("sps_search_get_ndt", self._h, C.byref(n))
        self.ndt = n.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.lib().sps_search_destroy(h)
            except Exception:
                pass
            self._h = None

    def run(self, d_in, d_dmt, d_cands, d_count):
        call("sps_search_run", self._h, _dptr(d_in), _dptr(d_dmt), _dptr(d_cands), _dptr(d_count), _stream())

    def set_schedule(self, depths):
        arr = (C.c_int32 * max(len(depths), 1))(*depths)
        call("sps_search_set_schedule", self._h, arr, len(depths))

    def set_staged(self, mode=1):
        """Staging mode of the merge passes (see FDMTGPU.set_staged)."""
        call("sps_search_set_staged", self._h, int(mode))

    def set_group_shape(self, lds_floats=8192, max_rows=16, tiles_per_block=8):
        call("sps_search_set_group_shape", self._h, int(lds_floats), int(max_rows), int(tiles_per_block))

    def get_schedule(self):
        n = C.c_int64()
        buf = (C.c_int32 * 64)()
        call("sps_search_get_schedule", self._h, buf, 64, C.byref(n))
        return list(buf)[: n.value]

    def enable_timing(self, on=True):
        call("sps_search_enable_timing", self._h, int(on))

    def stage_times_ms(self):
        """Per-stage durations of the last run (HIP events on the run's
        stream), in launch order: [statistics, mask transform, FDMT init,
        pass 1 .. pass P (each pass fuses the iterations its schedule entry
        names; the last carries the S/N division), separate S/N
        normalisation (0 when fused), boxcar]."""
        n = C.c_int64()
        buf = (C.c_double * 64)()
        call("sps_search_get_timing", self._h, buf, 64, C.byref(n))
        return list(buf)[: n.value]

    def stage_bytes(self):
        n = C.c_int64()
        buf = (C.c_double * 64)()
        call("sps_search_get_stage_bytes", self._h, buf, 64, C.byref(n))
        return list(buf)[: n.value]

    def sort_candidates(self, d_cands, d_count):
        """Sorts the run's candidates by (mdt, mt, mwidth) and returns their
        count.  Raises CapacityError (with .count) when the run found more
        than cand_capacity: the kept subset is then arbitrary and the caller
        re-gathers with gather() into a buffer of .count entries."""
        n = C.c_int64()
        st = _lib.lib().sps_search_sort_candidates(self._h, _dptr(d_cands), _dptr(d_count), C.byref(n), _stream())
        if st == _lib.SPS_ERR_CAPACITY:
            e = _lib.CapacityError(st, _lib.lib().sps_last_error().decode(errors="replace"))
            e.count = n.value
            raise e
        _lib.check(st)
        return n.value

    def gather(self, d_dmt, d_cands, d_count):
        """Re-runs only the boxcar gather of the last run over its S/N image
        (d_dmt as passed to run(), None when the run kept it internally) into
        d_cands ((cap, 4) int32), sorted; returns the count (CapacityError
        when it exceeds cap)."""
        n = C.c_int64()
        call("sps_search_gather", self._h, _dptr(d_dmt), _dptr(d_cands), int(d_cands.shape[0]), _dptr(d_count),
             C.byref(n), _stream())
        return n.value

    def cluster(self, d_dmt, wnd, d_bins):
        """The clustering leg of CChunk_gpu::process (Chunk_gpu.cu:398-425) on
        the last run's S/N image: all-widths gather (widths 1..wnd), three-key
        clustering with d_bins, delegates.  Returns (delegate candidates
        (numpy CAND_DTYPE, sorted by mt), group sizes, gathered count)."""
        nc, ng = C.c_int64(), C.c_int64()
        call("sps_search_cluster", self._h, _dptr(d_dmt), int(wnd), _dptr(d_bins), C.byref(nc), C.byref(ng),
             _stream())
        dele = np.zeros(max(ng.value, 1), CAND_DTYPE)
        gsz = np.zeros(max(ng.value, 1), np.int64)
        n = C.c_int64()
        call("sps_search_get_delegates", self._h, dele.ctypes.data_as(C.c_void_p),
             gsz.ctypes.data_as(C.POINTER(C.c_int64)), ng.value, C.byref(n))
        return dele[: ng.value], gsz[: ng.value], nc.value

    def write_log(self, filename, delt_bin, t_chunk_begin):
        """select_delegates::write_log of the last cluster() call."""
        call("sps_search_write_log", self._h, str(filename).encode(), float(delt_bin), float(t_chunk_begin))

    def candidates(self, d_dmt, d_cands, d_count):
        """Sorted candidates of the last run: (tensor (n, 4), n).  When the run
        overflowed d_cands, a buffer of exactly the count is allocated and
        re-gathered, so no candidate is ever dropped."""
        import torch

        try:
            n = self.sort_candidates(d_cands, d_count)
            return d_cands[:n], n
        except _lib.CapacityError as e:
            big = torch.empty((e.count, 4), dtype=torch.int32, device=d_cands.device)
            n = self.gather(d_dmt, big, d_count)
            return big[:n], n
