"""ctypes front for the CPU oracle (liboracle.so) and the reference build (_ref).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py; the product package never imports this module.

The coherent front-end (coherent_chirp, coherent_detect, coherent_chain) is
restated here in numpy.  It is pinned by what the mathematics fixes
(tests/test_oracle_coherent.py):
  * at DM 0 the chirp factor is the real taper / nbin;
  * the intensity pass conserves the kept intensity;
  * a pure tone lands in the row the index maps predict (derived
    independently in the test);
  * each row's dispersion shift is its rounded delay;
  * the CFdmt of a row-constant keep mask is causal and flat (the chunk
    pipeline's premise).
No reference test or fixture covers this path (cuFFT is not runnable here):
its FFT rounding against cuFFT is parity unpinned, bounded by the tolerance
against double-precision FFTs.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_i64 = C.c_int64
_f32 = C.c_float
_vp = C.c_void_p
_P = np.ctypeslib.ndpointer

CAND_DTYPE = np.dtype([("mt", np.int32), ("mdt", np.int32), ("mwidth", np.int32), ("msnr", np.float32)])


def build(ref=False):
    subprocess.check_call(["make", "-s", "-C", HERE], stdout=subprocess.DEVNULL)
    if ref:
        subprocess.check_call(["make", "-s", "-C", HERE, "ref"], stdout=subprocess.DEVNULL)


def _ptr(a):
    return a.ctypes.data_as(_vp) if a is not None else None


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.orc_last_error.restype = C.c_char_p
        L.orc_plan_create.restype = _vp
        L.orc_plan_create.argtypes = [_f32, _f32, _i64, _i64, _f32, _i64, _i64, _i64]
        L.orc_plan_destroy.argtypes = [_vp]
        for n in ("orc_plan_niters",):
            getattr(L, n).restype = _i64
            getattr(L, n).argtypes = [_vp]
        for n in ("orc_plan_nsub", "orc_plan_nrows", "orc_plan_nmap", "orc_plan_ncopy"):
            getattr(L, n).restype = _i64
            getattr(L, n).argtypes = [_vp, _i64]
        L.orc_plan_state_shape.argtypes = [_vp, _i64, _vp]
        L.orc_plan_grid.argtypes = [_vp, _i64, _vp, _vp]
        L.orc_plan_maps.argtypes = [_vp, _i64, _vp, _vp]
        L.orc_plan_df.argtypes = [_vp, _i64, C.POINTER(_f32), C.POINTER(_f32)]
        L.orc_fdmt_execute.argtypes = [_vp, _vp, _vp, C.c_int]
        L.orc_fdmt_initialise.argtypes = [_vp, _vp, _vp, C.c_int]
        L.orc_fdmt_bruteforce_f64.argtypes = [_vp, _vp, _vp]
        L.orc_fdmt_track_extent.argtypes = [_vp, _vp]
        L.orc_cfdmt_create.restype = _vp
        L.orc_cfdmt_create.argtypes = [_f32, _f32, C.c_int, C.c_int, C.c_int]
        L.orc_cfdmt_destroy.argtypes = [_vp]
        L.orc_cfdmt_niter.argtypes = [_vp]
        L.orc_cfdmt_quant.argtypes = [_vp, C.c_int]
        L.orc_cfdmt_config.argtypes = [_vp, C.c_int, _vp, _vp]
        L.orc_cfdmt_init_delta.argtypes = [_vp]
        L.orc_cfdmt_process.argtypes = [_vp, _vp, _vp, C.c_int]
        L.orc_stats.argtypes = [_vp, _vp, _i64, _i64, C.c_int, _vp, _vp, _vp, _vp]
        L.orc_normalise.argtypes = [_vp, _vp, _i64, _i64, C.c_int, _vp, _vp, _vp, _vp, _vp, _vp]
        L.orc_fdmt_normalization.argtypes = [_vp, _vp, _i64, _vp]
        L.orc_boxcar_pow2.restype = _i64
        L.orc_boxcar_pow2.argtypes = [_vp, _i64, _i64, _f32, C.c_int, _vp, _i64]
        L.orc_boxcar_pow2_alloc.restype = _vp
        L.orc_boxcar_pow2_alloc.argtypes = [_vp, _i64, _i64, _f32, C.c_int, C.POINTER(_i64)]
        L.orc_free.argtypes = [_vp]
        L.orc_boxcar_all.restype = _i64
        L.orc_boxcar_all.argtypes = [_vp, _i64, _i64, _f32, C.c_int, _vp, _i64]
        L.orc_cluster.restype = _i64
        L.orc_cluster.argtypes = [_vp, _i64, _vp, _vp, _vp, _vp, _vp]
        L.orc_detect_max.argtypes = [_vp, _vp, _i64, _i64, C.c_int, _vp, C.POINTER(_f32)]
        L.orc_select_delegates.argtypes = [_vp, _vp, _i64, _vp, _i64, _vp]
        L.orc_detect_tiles.argtypes = [C.c_int, _vp, _vp, _i64, _i64, C.c_int, C.c_int, _vp, C.POINTER(_f32)]
        L.orc_unpack_guppi.argtypes = [_vp, _i64, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _vp]
        L.orc_unpack_lofar.argtypes = [_vp, _i64, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _vp]
        L.orc_num_threads.restype = C.c_int
        L.orc_set_num_threads.argtypes = [C.c_int]
        _lib = L
    return _lib


def err():
    return lib().orc_last_error().decode()


class Plan:
    """FDMT plan (fdmt_base.cpp:37-53,119-279) as numpy tables."""

    def __init__(self, fmin, fmax, nchans, nsamps, tsamp=1.0, dt_max=0, dt_step=1, dt_min=0):
        L = lib()
        self.h = L.orc_plan_create(fmin, fmax, nchans, nsamps, tsamp, dt_max, dt_step, dt_min)
        if not self.h:
            raise RuntimeError(err())
        self.nchans, self.nsamps = nchans, nsamps
        self.niters = L.orc_plan_niters(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_plan_destroy(self.h)
            self.h = None

    def nrows(self, i):
        return lib().orc_plan_nrows(self.h, i)

    def state_shape(self, i):
        out = np.zeros(5, np.int64)
        lib().orc_plan_state_shape(self.h, i, _ptr(out))
        return out

    def grid(self, i):
        ns = lib().orc_plan_nsub(self.h, i)
        off = np.zeros(ns + 1, np.int64)
        lib().orc_plan_grid(self.h, i, _ptr(off), None)
        g = np.zeros(max(int(off[-1]), 1), np.int64)
        lib().orc_plan_grid(self.h, i, _ptr(off), _ptr(g))
        return off, g[: off[-1]]

    def maps(self, i):
        nm = lib().orc_plan_nmap(self.h, i)
        nc = lib().orc_plan_ncopy(self.h, i)
        m = np.zeros((max(nm, 1), 7), np.int64)
        c = np.zeros((max(nc, 1), 4), np.int64)
        lib().orc_plan_maps(self.h, i, _ptr(m), _ptr(c))
        return m[:nm], c[:nc]

    def df(self, i):
        a, b = _f32(), _f32()
        lib().orc_plan_df(self.h, i, C.byref(a), C.byref(b))
        return a.value, b.value

    def execute(self, x, init_mode=0):
        x = np.ascontiguousarray(x, np.float32)
        assert x.size == self.nchans * self.nsamps
        out = np.zeros((self.nrows(self.niters), self.nsamps), np.float32)
        if lib().orc_fdmt_execute(self.h, _ptr(x), _ptr(out), init_mode) != 0:
            raise RuntimeError(err())
        return out

    def initialise(self, x, init_mode=0):
        x = np.ascontiguousarray(x, np.float32)
        out = np.zeros((self.nrows(0), self.nsamps), np.float32)
        if lib().orc_fdmt_initialise(self.h, _ptr(x), _ptr(out), init_mode) != 0:
            raise RuntimeError(err())
        return out

    def bruteforce(self, x):
        x = np.ascontiguousarray(x, np.float32)
        out = np.zeros((self.nrows(self.niters), self.nsamps), np.float64)
        if lib().orc_fdmt_bruteforce_f64(self.h, _ptr(x), _ptr(out)) != 0:
            raise RuntimeError(err())
        return out

    def track_extent(self):
        out = np.zeros(self.nrows(self.niters), np.int64)
        lib().orc_fdmt_track_extent(self.h, _ptr(out))
        return out


class CFdmt:
    """Legacy CFdmtCpu (FdmtCpu.cpp:131-249)."""

    def __init__(self, fmin, fmax, nchan, cols, imaxdt):
        self.h = lib().orc_cfdmt_create(fmin, fmax, nchan, cols, imaxdt)
        if not self.h:
            raise RuntimeError(err())
        self.nchan, self.cols, self.imaxdt = nchan, cols, imaxdt
        self.niter = lib().orc_cfdmt_niter(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_cfdmt_destroy(self.h)
            self.h = None

    def config(self, i):
        q = lib().orc_cfdmt_quant(self.h, i)
        cs = np.zeros(q + 1, np.int32)
        fr = np.zeros(q + 1, np.float32)
        lib().orc_cfdmt_config(self.h, i, _ptr(cs), _ptr(fr))
        return cs, fr

    def init_delta(self):
        return lib().orc_cfdmt_init_delta(self.h)

    def process(self, x, b_ones=False):
        out = np.zeros((self.imaxdt, self.cols), np.float32)
        xa = None if x is None else np.ascontiguousarray(x, np.float32)
        if lib().orc_cfdmt_process(self.h, _ptr(xa), _ptr(out), int(b_ones)) != 0:
            raise RuntimeError(err())
        return out


def stats(x, mode=1):
    """Row/global statistics and keep flags (Statistical_preprocessing.cu:15-88)."""
    R, N = x.shape
    rm = np.zeros(R, np.float64)
    rv = np.zeros(R, np.float64)
    keep = np.zeros(R, np.uint8)
    st = np.zeros(5, np.float64)
    if x.dtype == np.uint8:
        xa = np.ascontiguousarray(x)
        lib().orc_stats(None, _ptr(xa), R, N, mode, _ptr(rm), _ptr(rv), _ptr(keep), _ptr(st))
    else:
        xa = np.ascontiguousarray(x, np.float32)
        lib().orc_stats(_ptr(xa), None, R, N, mode, _ptr(rm), _ptr(rv), _ptr(keep), _ptr(st))
    return rm, rv, keep, st


def normalise(x, mode=1):
    """calc_fdmt_inputs: returns (normalised fp32 image, mask, stats tuple)."""
    R, N = x.shape
    rm, rv, keep, st = stats(x, mode)
    out = np.zeros((R, N), np.float32)
    mask = np.zeros((R, N), np.float32)
    xa = np.ascontiguousarray(x)
    if x.dtype == np.uint8:
        lib().orc_normalise(None, _ptr(xa), R, N, mode, _ptr(rm), _ptr(rv), _ptr(keep), _ptr(st), _ptr(out), _ptr(mask))
    else:
        xa = np.ascontiguousarray(x, np.float32)
        lib().orc_normalise(_ptr(xa), None, R, N, mode, _ptr(rm), _ptr(rv), _ptr(keep), _ptr(st), _ptr(out), _ptr(mask))
    return out, mask, (rm, rv, keep, st)


def normalise_with_stats(x, mode, st):
    """calc_fdmt_inputs' normalisation of x with statistics st = (row means,
    row variances, keep flags, stats) computed elsewhere (for instance over
    the whole of a beam of which x is a leading time slice)."""
    rm, rv, keep, stv = st
    R, N = x.shape
    out = np.zeros((R, N), np.float32)
    mask = np.zeros((R, N), np.float32)
    if x.dtype == np.uint8:
        xa = np.ascontiguousarray(x)
        lib().orc_normalise(None, _ptr(xa), R, N, mode, _ptr(rm), _ptr(rv), _ptr(keep), _ptr(stv), _ptr(out),
                            _ptr(mask))
    else:
        xa = np.ascontiguousarray(x, np.float32)
        lib().orc_normalise(_ptr(xa), None, R, N, mode, _ptr(rm), _ptr(rv), _ptr(keep), _ptr(stv), _ptr(out),
                            _ptr(mask))
    return out, mask


def fdmt_normalization(a, n):
    a = np.ascontiguousarray(a, np.float32)
    n = np.ascontiguousarray(n, np.float32)
    out = np.zeros_like(a)
    lib().orc_fdmt_normalization(_ptr(a), _ptr(n), a.size, _ptr(out))
    return out


def boxcar_pow2(img, thr, q):
    """Power-of-two boxcar candidates (widths 1..2^q) in (mdt, mt, mwidth) order."""
    img = np.ascontiguousarray(img, np.float32)
    n = _i64()
    p = lib().orc_boxcar_pow2_alloc(_ptr(img), img.shape[0], img.shape[1], thr, q, C.byref(n))
    out = np.zeros(n.value, CAND_DTYPE)
    if n.value:
        C.memmove(out.ctypes.data, p, out.nbytes)
        lib().orc_free(p)
    return out


def boxcar_all(img, thr, w, cap=1 << 20):
    img = np.ascontiguousarray(img, np.float32)
    out = np.zeros(cap, CAND_DTYPE)
    n = lib().orc_boxcar_all(_ptr(img), img.shape[0], img.shape[1], thr, w, _ptr(out), cap)
    if n > cap:
        return boxcar_all(img, thr, w, cap=int(n))
    return out[:n].copy()


def cluster(cands, bins):
    """Returns (order, group_off, delegates (sorted by mt), group_size)."""
    c = np.ascontiguousarray(cands, CAND_DTYPE)
    n = c.size
    b = np.asarray(bins, np.int32)
    order = np.zeros(max(n, 1), np.int32)
    goff = np.zeros(n + 1, np.int64)
    dele = np.zeros(max(n, 1), np.int32)
    gsz = np.zeros(max(n, 1), np.int64)
    ng = lib().orc_cluster(_ptr(c), n, _ptr(b), _ptr(order), _ptr(goff), _ptr(dele), _ptr(gsz))
    return order[:n], goff[: ng + 1], dele[:ng], gsz[:ng]


def detect_max(a, norm, w):
    a = np.ascontiguousarray(a, np.float32)
    na = None if norm is None else np.ascontiguousarray(norm, np.float32)
    o = np.zeros(3, np.int32)
    s = _f32()
    rc = lib().orc_detect_max(_ptr(a), _ptr(na), a.shape[0], a.shape[1], w, _ptr(o), C.byref(s))
    if rc != 0:
        return None
    return int(o[0]), int(o[1]), int(o[2]), s.value


def select_delegates(cands, indices, group_begin):
    """Delegates of caller groups (select_delegates.cu:16-51)."""
    c = np.ascontiguousarray(cands, CAND_DTYPE)
    idx = np.ascontiguousarray(indices, np.int32)
    b = np.ascontiguousarray(group_begin, np.int32)
    out = np.zeros(max(b.size, 1), np.int32)
    lib().orc_select_delegates(_ptr(c), _ptr(idx), idx.size, _ptr(b), b.size, _ptr(out))
    return out[: b.size]


def unpack(layout, raw, nsamp, noverlap, nbin, nfft, npol, nchan):
    """8-bit raw voltages -> complex64 blocks [pol][fft][chan][bin]
    (unpack_guppi/kernel.cu:81-122 LOFAR, :161-187 GUPPI)."""
    x = np.ascontiguousarray(raw, np.int8).reshape(-1)
    if x.size < nsamp * nchan * npol:
        raise ValueError("raw chunk too short")
    out = np.zeros((npol // 2, nfft, nchan, nbin), np.complex64)
    f = lib().orc_unpack_guppi if layout == "guppi" else lib().orc_unpack_lofar
    f(_ptr(x), nsamp, noverlap, nbin, nfft, npol, nchan, _ptr(out))
    return out


def coherent_chirp(spec, dm, f_min, f_max, nchan, len_sft, nbin, nfft, npol):
    """kernel_el_wise_mult_onthe_fly (CDMT/Chunk_fly_gpu.cu:94-164) on
    spectra [pol][fft][chan][bin]: bin col of channel c (col = (isft * mbin +
    ibin + nbin/2) mod nbin) times the factor of sub-band isft, bin ibin.
    Restated in numpy: the sub-band centre rounded to float, the rest in
    double, the complex product in float32 without contraction."""
    mbin = nbin // len_sft
    bw = float(np.float32(f_max) - np.float32(f_min))  # a float difference
    bw_sub = bw / nchan
    bw_chan = bw_sub / len_sft
    ichan = np.arange(nchan, dtype=np.float64)[:, None, None]
    isft = np.arange(len_sft, dtype=np.float64)[None, :, None]
    ibin = np.arange(mbin, dtype=np.float64)[None, None, :]
    f0 = (float(np.float32(f_min)) + bw_sub * (0.5 + ichan) + bw_chan * (isft - len_sft / 2.0 + 0.5))
    f0 = f0.astype(np.float32).astype(np.float64)
    bf = -0.5 * bw_chan + (ibin + 0.5) * bw_chan / mbin
    taper = 1.0 / np.sqrt(1.0 + np.power(bf / (0.47 * bw_chan), 80.0))
    r = bf / f0
    delay = dm * r * r / (f0 + bf) * 4.148808e9
    t = -np.fmod(delay, 1.0) * 2.0
    c, sn = np.cos(np.pi * t), np.sin(np.pi * t)
    dx = (c * taper / float(nbin)).astype(np.float32)
    dy = (sn * taper / float(nbin)).astype(np.float32)
    col = ((np.arange(len_sft)[:, None] * mbin + np.arange(mbin)[None, :] + nbin // 2) % nbin)
    fx = np.zeros((nchan, nbin), np.float32)
    fy = np.zeros((nchan, nbin), np.float32)
    fx[:, col] = dx.reshape(nchan, len_sft, mbin)
    fy[:, col] = dy.reshape(nchan, len_sft, mbin)
    a = np.asarray(spec, np.complex64).reshape(npol // 2, nfft, nchan, nbin)
    ax, ay = a.real.astype(np.float32), a.imag.astype(np.float32)
    ox = ax * fx - ay * fy
    oy = ax * fy + ay * fx
    return (ox + 1j * oy).astype(np.complex64)


def coherent_detect(buf, dm, f_min, f_max, nchan, len_sft, nbin, nfft, npol, noverlap, tsamp):
    """The three reference passes after the inverse FFT, literally
    (CDMT/Chunk_gpu.cu:290-322): roll_rows_normalize_sum_kernel (:476-493),
    transpose_unpadd_intensity (:497-515), dedisperse (:613-632).  Returns
    the float32 waterfall [nchan * len_sft][msamp]."""
    mbin = nbin // len_sft
    nov = noverlap // len_sft
    madj = mbin - 2 * nov
    msamp = nfft * madj
    a = np.asarray(buf, np.complex64).reshape(npol // 2, nfft * nchan, nbin)
    # roll, normalise, sum: rolled[row][(i + nbin/2) % nbin] += float((x^2 + y^2)/nbin/nbin)
    rolled = np.zeros((nfft * nchan, nbin), np.float32)
    dst = (np.arange(nbin) + nbin // 2) % nbin
    for p in range(npol // 2):
        x = a[p].real.astype(np.float64)
        y = a[p].imag.astype(np.float64)
        rolled[:, dst] = rolled[:, dst] + ((x * x + y * y) / nbin / nbin).astype(np.float32)
    # transpose and unpad (nsub = nchan coarse channels, len_sft sub-channels)
    r4 = rolled.reshape(nfft, nchan, len_sft, mbin)
    inten = np.zeros((nchan * len_sft, msamp), np.float32)
    for isub in range(nchan):
        for ich in range(len_sft):
            row = isub * len_sft + len_sft - ich - 1
            seg = r4[:, nchan - isub - 1, ich, nov:nov + madj]  # [fft][madj]
            inten[row] = seg.reshape(-1)
    # dedisperse
    foff = (float(np.float32(f_max)) - float(np.float32(f_min))) / (nchan * len_sft)
    fmin = float(np.float32(f_min))
    tsw = float(np.float32(np.float32(len_sft) * np.float32(tsamp)))
    out = np.zeros_like(inten)
    nrows = nchan * len_sft
    for row in range(nrows):
        fr = (nrows - 1 - row) * foff + fmin
        d = 4.148808e3 * dm * (1.0 / (fmin * fmin) - 1.0 / (fr * fr))
        q = d / tsw
        ish = int(np.floor(q + 0.5)) if q >= 0 else -int(np.floor(-q + 0.5))  # C round(): halves away from 0
        out[row, (np.arange(msamp) + ish) % msamp] = inten[row]
    return out


def coherent_chain(spec_time, dm, f_min, f_max, nchan, len_sft, nbin, nfft, npol, noverlap, tsamp):
    """The whole front-end in double-precision FFTs (numpy): forward FFT of
    the blocks, the chirp, the unnormalised inverse FFT of every sub-band
    and the three passes above; a tolerance reference for the rocFFT path."""
    mbin = nbin // len_sft
    a = np.asarray(spec_time, np.complex64).reshape(npol // 2, nfft, nchan, nbin)
    f = np.fft.fft(a.astype(np.complex128), axis=-1).astype(np.complex64)
    ch = coherent_chirp(f, dm, f_min, f_max, nchan, len_sft, nbin, nfft, npol)
    sub = ch.reshape(npol // 2, nfft, nchan, len_sft, mbin).astype(np.complex128)
    inv = (np.fft.ifft(sub, axis=-1) * mbin).astype(np.complex64).reshape(npol // 2, nfft, nchan, nbin)
    return coherent_detect(inv, dm, f_min, f_max, nchan, len_sft, nbin, nfft, npol, noverlap, tsamp)


def detect_tiles(mode, a, norm, w, tile):
    """The tiled detectors of doDetectSignal_v1 (mode 1) / _v2 (mode 2):
    (width, row, col, value) or None."""
    a = np.ascontiguousarray(a, np.float32)
    na = None if norm is None else np.ascontiguousarray(norm, np.float32)
    o = np.zeros(3, np.int32)
    s = _f32()
    rc = lib().orc_detect_tiles(int(mode), _ptr(a), _ptr(na), a.shape[0], a.shape[1], int(w), int(tile), _ptr(o),
                                C.byref(s))
    if rc != 0:
        return None
    return int(o[0]), int(o[1]), int(o[2]), s.value


# --------------------------------------------------------------------------
# reference build (oracle/_ref), present only where /root/reference was
# available at build time
# --------------------------------------------------------------------------
REF_DIR = os.path.join(HERE, "_ref")


def ref_available():
    return all(os.path.exists(os.path.join(REF_DIR, f"libref_{k}.so")) for k in ("pravir", "restrict", "cfdmt"))


_ref = {}


def ref_lib(kind):
    if kind not in _ref:
        L = C.CDLL(os.path.join(REF_DIR, f"libref_{kind}.so"))
        if kind == "cfdmt":
            L.ref_cfdmt_create.restype = _vp
            L.ref_cfdmt_create.argtypes = [_f32, _f32, C.c_int, C.c_int, C.c_int]
            L.ref_cfdmt_destroy.argtypes = [_vp]
            L.ref_cfdmt_niter.argtypes = [_vp]
            L.ref_cfdmt_quant.argtypes = [_vp, C.c_int]
            L.ref_cfdmt_config.argtypes = [_vp, C.c_int, _vp, _vp]
            L.ref_cfdmt_process.argtypes = [_vp, _vp, _vp, C.c_int]
        else:
            L.ref_fdmt_create.restype = _vp
            L.ref_fdmt_create.argtypes = [_f32, _f32, _i64, _i64, _f32, _i64, _i64, _i64]
            L.ref_fdmt_destroy.argtypes = [_vp]
            L.ref_fdmt_niters.restype = _i64
            L.ref_fdmt_niters.argtypes = [_vp]
            for n in ("ref_fdmt_nsub", "ref_fdmt_nmap", "ref_fdmt_ncopy"):
                getattr(L, n).restype = _i64
                getattr(L, n).argtypes = [_vp, _i64]
            L.ref_fdmt_state_shape.argtypes = [_vp, _i64, _vp]
            L.ref_fdmt_grid.argtypes = [_vp, _i64, _vp, _vp]
            L.ref_fdmt_maps.argtypes = [_vp, _i64, _vp, _vp]
            L.ref_fdmt_df.argtypes = [_vp, _i64, C.POINTER(_f32), C.POINTER(_f32)]
            L.ref_fdmt_execute.argtypes = [_vp, _vp, _i64, _vp, _i64]
            L.ref_fdmt_dm_grid.argtypes = [_vp, _vp]
            L.ref_fdmt_set_num_threads.argtypes = [C.c_int]
        _ref[kind] = L
    return _ref[kind]


class RefFDMT:
    """The reference FDMTCPU (FDMT_PRAVIR or FDMT_PRAVIR_RESTRICT) via _ref."""

    def __init__(self, kind, fmin, fmax, nchans, nsamps, tsamp=1.0, dt_max=0, dt_step=1, dt_min=0):
        self.L = ref_lib(kind)
        self.h = self.L.ref_fdmt_create(fmin, fmax, nchans, nsamps, tsamp, dt_max, dt_step, dt_min)
        if not self.h:
            raise RuntimeError("reference plan builder threw")
        self.nchans, self.nsamps = nchans, nsamps
        self.niters = self.L.ref_fdmt_niters(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.ref_fdmt_destroy(self.h)
            self.h = None

    def state_shape(self, i):
        out = np.zeros(5, np.int64)
        self.L.ref_fdmt_state_shape(self.h, i, _ptr(out))
        return out

    def grid(self, i):
        ns = self.L.ref_fdmt_nsub(self.h, i)
        off = np.zeros(ns + 1, np.int64)
        total = int(self.state_shape(i)[3])
        g = np.zeros(max(total, 1), np.int64)
        self.L.ref_fdmt_grid(self.h, i, _ptr(off), _ptr(g))
        return off, g[: off[-1]]

    def maps(self, i):
        nm = self.L.ref_fdmt_nmap(self.h, i)
        nc = self.L.ref_fdmt_ncopy(self.h, i)
        m = np.zeros((max(nm, 1), 7), np.int64)
        c = np.zeros((max(nc, 1), 4), np.int64)
        self.L.ref_fdmt_maps(self.h, i, _ptr(m), _ptr(c))
        return m[:nm], c[:nc]

    def df(self, i):
        a, b = _f32(), _f32()
        self.L.ref_fdmt_df(self.h, i, C.byref(a), C.byref(b))
        return a.value, b.value

    def execute(self, x):
        x = np.ascontiguousarray(x, np.float32)
        rows = int(self.state_shape(self.niters)[3])
        out = np.zeros((rows, self.nsamps), np.float32)
        if self.L.ref_fdmt_execute(self.h, _ptr(x), x.size, _ptr(out), out.size) != 0:
            raise RuntimeError("reference execute threw")
        return out

    def dm_grid(self):
        rows = int(self.state_shape(self.niters)[3])
        out = np.zeros(rows, np.float32)
        self.L.ref_fdmt_dm_grid(self.h, _ptr(out))
        return out


class RefCFdmt:
    """The reference CFdmtCpu (FDMT/FdmtCpu.cpp) via _ref."""

    def __init__(self, fmin, fmax, nchan, cols, imaxdt):
        self.L = ref_lib("cfdmt")
        self.h = self.L.ref_cfdmt_create(fmin, fmax, nchan, cols, imaxdt)
        self.nchan, self.cols, self.imaxdt = nchan, cols, imaxdt
        self.niter = self.L.ref_cfdmt_niter(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.ref_cfdmt_destroy(self.h)
            self.h = None

    def config(self, i):
        q = self.L.ref_cfdmt_quant(self.h, i)
        cs = np.zeros(q + 1, np.int32)
        fr = np.zeros(q + 1, np.float32)
        self.L.ref_cfdmt_config(self.h, i, _ptr(cs), _ptr(fr))
        return cs, fr

    def process(self, x, b_ones=False):
        out = np.zeros((self.imaxdt, self.cols), np.float32)
        xa = None if x is None else np.ascontiguousarray(x, np.float32)
        self.L.ref_cfdmt_process(self.h, _ptr(xa), _ptr(out), int(b_ones))
        return out
