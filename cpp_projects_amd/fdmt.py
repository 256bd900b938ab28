"""Python mirror of the reference FDMT classes over the C ABI.

FDMTGPU mirrors FDMT/FDMTGPU (FDMT_PRAVIR_RESTRICT/fdmt_base.hpp:37-85,
fdmt_gpu.cuh:68-89) and FDMTCPU (fdmt_cpu.hpp:5-20); CFdmtGpu and CFdmtCpu
mirror the legacy classes (FDMT/FdmtGpu.cuh:5-57, FdmtCpu.h:6-34).  The GPU
classes take torch tensors on the HIP device and run on torch's current
stream; the CPU classes take numpy arrays and return when the result is in
host memory (the transform runs on the GPU in both).  Plan queries need no
GPU.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import SPS_F32, SPS_INIT_CFDMT, SPS_INIT_MEAN, SPS_INIT_RESTRICT, SPS_U8, call  # noqa: F401


def _stream():
    import torch

    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dptr(t):
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("expected a device tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return C.c_void_p(t.data_ptr())


def _dtype_code(t):
    import torch

    if t.dtype == torch.float32:
        return SPS_F32
    if t.dtype == torch.uint8:
        return SPS_U8
    raise ValueError(f"unsupported waterfall dtype {t.dtype}")


class _Handle:
    def __init__(self):
        self._h = C.c_void_p()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.lib().sps_fdmt_destroy(h)
            except Exception:
                pass
            self._h = None

    # ---- queries (host only) ----
    def get_niters(self):
        n = C.c_int64()
        call("sps_fdmt_get_niters", self._h, C.byref(n))
        return n.value

    def state_shape(self, i):
        s = (C.c_int64 * 5)()
        call("sps_fdmt_get_state_shape", self._h, i, s)
        return np.array(list(s), np.int64)

    def output_shape(self):
        r, c = C.c_int64(), C.c_int64()
        call("sps_fdmt_get_output_shape", self._h, C.byref(r), C.byref(c))
        return r.value, c.value

    def rows(self, i):
        n = C.c_int64()
        call("sps_fdmt_get_rows", self._h, i, C.byref(n), None)
        out = np.zeros((max(n.value, 1), 4), np.int64)
        call("sps_fdmt_get_rows", self._h, i, C.byref(n), out.ctypes.data_as(C.POINTER(C.c_int64)))
        return out[: n.value]

    def set_schedule(self, depths):
        """Merge passes: each fuses 1..3 consecutive iterations, 1..4 when
        staged (sum = niters)."""
        arr = (C.c_int32 * max(len(depths), 1))(*depths)
        call("sps_fdmt_set_schedule", self._h, arr, len(depths))

    def set_staged(self, mode=1):
        """0: tree passes; 1: LDS-staged passes (grouped output rows, each
        input row read once per time tile), the final pass on the tree kernel
        when its depth allows; 2: as 1 with the mean init fused into the first
        pass; 3, 4: as 1, 2 with the final pass staged too."""
        call("sps_fdmt_set_staged", self._h, int(mode))

    def set_group_shape(self, lds_floats=8192, max_rows=16, tiles_per_block=8):
        call("sps_fdmt_set_group_shape", self._h, int(lds_floats), int(max_rows), int(tiles_per_block))

    def get_schedule(self):
        n = C.c_int64()
        buf = (C.c_int32 * 64)()
        call("sps_fdmt_get_schedule", self._h, buf, 64, C.byref(n))
        return list(buf)[: n.value]

    # ---- device work ----
    def _execute(self, waterfall, waterfall_size, dmt, dmt_size, ones=False):
        code = SPS_F32 if waterfall is None else _dtype_code(waterfall)
        call("sps_fdmt_execute", self._h, _dptr(waterfall), code, int(waterfall_size), _dptr(dmt), int(dmt_size),
             int(bool(ones)), _stream())

    def adjoint(self, grad_dmt, grad_waterfall):
        call("sps_fdmt_adjoint", self._h, _dptr(grad_dmt), grad_dmt.numel(), _dptr(grad_waterfall),
             grad_waterfall.numel(), _stream())


class FDMTGPU(_Handle):
    """FDMT(f_min, f_max, nchans, nsamps, tsamp, dt_max, dt_step=1, dt_min=0)
    (fdmt_base.hpp:39-41).  init_mode selects the FDMT_PRAVIR running mean
    (default) or the FDMT_PRAVIR_RESTRICT form (SPS_INIT_RESTRICT)."""

    def __init__(self, f_min, f_max, nchans, nsamps, tsamp, dt_max, dt_step=1, dt_min=0, init_mode=SPS_INIT_MEAN):
        super().__init__()
        call("sps_fdmt_create", float(f_min), float(f_max), int(nchans), int(nsamps), float(tsamp), int(dt_max),
             int(dt_step), int(dt_min), int(init_mode), C.byref(self._h))
        self.nchans, self.nsamps = int(nchans), int(nsamps)

    def get_df(self):
        df, corr = C.c_float(), C.c_float()
        call("sps_fdmt_get_df", self._h, C.byref(df), C.byref(corr))
        return df.value

    def get_correction(self):
        df, corr = C.c_float(), C.c_float()
        call("sps_fdmt_get_df", self._h, C.byref(df), C.byref(corr))
        return corr.value

    def get_m_nsamps(self):
        return self.nsamps

    def get_dt_grid_final(self):
        n = C.c_int64()
        call("sps_fdmt_get_dt_grid_final", self._h, None, 0, C.byref(n))
        out = np.zeros(max(n.value, 1), np.int64)
        call("sps_fdmt_get_dt_grid_final", self._h, out.ctypes.data_as(C.POINTER(C.c_int64)), n.value, C.byref(n))
        return out[: n.value]

    def get_dm_grid_final(self):
        n = C.c_int64()
        call("sps_fdmt_get_dm_grid_final", self._h, None, 0, C.byref(n))
        out = np.zeros(max(n.value, 1), np.float32)
        call("sps_fdmt_get_dm_grid_final", self._h, out.ctypes.data_as(C.POINTER(C.c_float)), n.value, C.byref(n))
        return out[: n.value]

    # plan tables (FDMTPlan, fdmt_base.hpp:20-35), iteration i
    def grid(self, i):
        ns = C.c_int64()
        P = C.POINTER(C.c_int64)
        call("sps_fdmt_get_plan_iter", self._h, i, C.byref(ns), None, None, None, None, None, None, None, None)
        off = np.zeros(ns.value + 1, np.int64)
        call("sps_fdmt_get_plan_iter", self._h, i, C.byref(ns), off.ctypes.data_as(P), None, None, None, None, None,
             None, None)
        g = np.zeros(max(int(off[-1]), 1), np.int64)
        call("sps_fdmt_get_plan_iter", self._h, i, None, None, g.ctypes.data_as(P), None, None, None, None, None, None)
        return off, g[: off[-1]]

    def maps(self, i):
        P = C.POINTER(C.c_int64)
        nm, nc = C.c_int64(), C.c_int64()
        call("sps_fdmt_get_plan_iter", self._h, i, None, None, None, C.byref(nm), None, C.byref(nc), None, None, None)
        m = np.zeros((max(nm.value, 1), 7), np.int64)
        c = np.zeros((max(nc.value, 1), 4), np.int64)
        call("sps_fdmt_get_plan_iter", self._h, i, None, None, None, None, m.ctypes.data_as(P), None,
             c.ctypes.data_as(P), None, None)
        return m[: nm.value], c[: nc.value]

    def df(self, i):
        a, b = C.c_float(), C.c_float()
        call("sps_fdmt_get_plan_iter", self._h, i, None, None, None, None, None, None, None, C.byref(a), C.byref(b))
        return a.value, b.value

    @property
    def niters(self):
        return self.get_niters()

    def execute(self, waterfall, waterfall_size, dmt, dmt_size):
        """FDMT::execute (fdmt_base.hpp:55-56): device waterfall (f32 or u8)
        of nchans*nsamps, device dmt of ndt*nsamps."""
        self._execute(waterfall, waterfall_size, dmt, dmt_size)

    def initialise(self, waterfall, state):
        call("sps_fdmt_initialise", self._h, _dptr(waterfall), _dtype_code(waterfall), _dptr(state), state.numel(),
             _stream())


class CFdmtGpu(_Handle):
    """CFdmtGpu(Fmin, Fmax, nchan, cols, imaxDt) (FdmtGpu.cuh:12-18)."""

    def __init__(self, f_min, f_max, nchan, cols, imaxdt):
        super().__init__()
        call("sps_cfdmt_create", float(f_min), float(f_max), int(nchan), int(cols), int(imaxdt), C.byref(self._h))
        self.m_nchan, self.m_cols, self.m_imaxDt = int(nchan), int(cols), int(imaxdt)

    def process_image(self, d_in, d_out, b_ones):
        """CFdmtGpu::process_image (FdmtGpu.cuh:50-53); d_in may be None when b_ones."""
        n_in = self.m_nchan * self.m_cols
        self._execute(d_in, n_in, d_out, d_out.numel(), ones=b_ones)


def _host_f32(a, n, what):
    if a is None:
        return None
    if not isinstance(a, np.ndarray) or a.dtype != np.float32 or not a.flags["C_CONTIGUOUS"]:
        raise ValueError(f"{what} must be a C-contiguous float32 numpy array")
    if a.size != n:
        raise _lib.InvalidArgument(_lib.SPS_ERR_INVALID_ARGUMENT, f"Invalid size of {what}")
    return a.ctypes.data_as(C.c_void_p)


class FDMTCPU(FDMTGPU):
    """FDMTCPU(f_min, f_max, nchans, nsamps, tsamp, dt_max, dt_step=1, dt_min=0)
    (fdmt_cpu.hpp:5-20): execute / initialise take host (numpy float32)
    arrays.  set_num_threads records the thread count a caller asks for; no
    host computation here uses it."""

    _threads = 0

    @staticmethod
    def set_num_threads(nthreads):
        FDMTCPU._threads = int(nthreads)

    def execute(self, waterfall, waterfall_size, dmt, dmt_size):
        if int(waterfall_size) != self.nchans * self.nsamps:
            raise _lib.InvalidArgument(_lib.SPS_ERR_INVALID_ARGUMENT, "Invalid size of waterfall")
        call("sps_fdmt_execute_host", self._h, _host_f32(waterfall, int(waterfall_size), "waterfall"),
             int(waterfall_size), _host_f32(dmt, int(dmt_size), "dmt"), int(dmt_size), 0)

    def initialise(self, waterfall, state):
        call("sps_fdmt_initialise_host", self._h, _host_f32(waterfall, self.nchans * self.nsamps, "waterfall"),
             _host_f32(state, state.size, "state"), state.size)


class CFdmtCpu(CFdmtGpu):
    """CFdmtCpu(Fmin, Fmax, nchan, cols, imaxDt) (FdmtCpu.h:6-34): host
    (numpy float32) input and output; m_iNumIter, m_parrQuantMtrx_h,
    m_pparrRowsCumSum_h and m_pparrFreq_h hold the legacy configuration."""

    def __init__(self, f_min, f_max, nchan, cols, imaxdt):
        super().__init__(f_min, f_max, nchan, cols, imaxdt)
        self.m_iNumIter = self.get_niters()
        self.m_parrQuantMtrx_h, self.m_pparrRowsCumSum_h, self.m_pparrFreq_h = [], [], []
        for i in range(self.m_iNumIter + 1):
            q = C.c_int32()
            call("sps_cfdmt_get_config", self._h, i, C.byref(q), None, None)
            cs = np.zeros(q.value + 1, np.int32)
            fr = np.zeros(q.value + 1, np.float32)
            call("sps_cfdmt_get_config", self._h, i, C.byref(q), cs.ctypes.data_as(C.POINTER(C.c_int32)),
                 fr.ctypes.data_as(C.POINTER(C.c_float)))
            self.m_parrQuantMtrx_h.append(q.value)
            self.m_pparrRowsCumSum_h.append(cs)
            self.m_pparrFreq_h.append(fr)

    def process_image(self, h_in, h_out, b_ones):
        """CFdmtCpu::process_image (FdmtCpu.cpp:197-249); h_in may be None when b_ones."""
        n_in = self.m_nchan * self.m_cols
        call("sps_fdmt_execute_host", self._h, None if b_ones and h_in is None else _host_f32(h_in, n_in, "input"),
             n_in, _host_f32(h_out, self.m_imaxDt * self.m_cols, "output"), self.m_imaxDt * self.m_cols,
             int(bool(b_ones)))
