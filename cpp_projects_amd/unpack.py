"""Python mirror of the sessions' 8-bit raw-voltage unpack over the C ABI
(CSession_guppi_gpu::unpack_chunk, CDMT/Session_guppi_gpu.cu:53-60, and
CSession_lofar_gpu::unpack_chunk, CDMT/Session_lofar_gpu.cu:54-67).  Device
buffers are torch tensors: the input int8, the output complex64 (or float32
re/im pairs) of nfft * nchan * nbin * npol / 2 values, laid out
[pol][fft][chan][bin].
"""
from ._lib import call
from .fdmt import _dptr, _stream


def _check(d_in, d_out, n_in, n_out):
    if d_in.numel() < n_in:
        raise ValueError(f"input holds {d_in.numel()} bytes, needs {n_in}")
    nf = d_out.numel() * (2 if d_out.is_complex() else 1)
    if nf < 2 * n_out:
        raise ValueError(f"output holds {nf} floats, needs {2 * n_out}")


def out_size(nbin, nfft, npol, nchan):
    """Complex values of one unpacked chunk."""
    return nfft * nchan * nbin * (npol // 2)


def unpack_chunk_guppi(d_in, nsamp, noverlap, nbin, nfft, npol, nchan, d_out):
    """GUPPI: d_in[chan][nsamp][npol] signed bytes (unpack_chunk_guppi_gpu,
    CDMT/Session_guppi_gpu.cu:116-140)."""
    _check(d_in, d_out, nchan * nsamp * npol, out_size(nbin, nfft, npol, nchan))
    call("sps_unpack_guppi", _dptr(d_in), int(nsamp), int(noverlap), int(nbin), int(nfft), int(npol), int(nchan),
         _dptr(d_out), _stream())


def unpack_chunk_lofar(d_in, nsamp, noverlap, nbin, nfft, npol, nchan, d_out):
    """LOFAR: d_in[npol][nsamp][nchan] signed bytes, components 2p / 2p + 1
    the real / imaginary parts of pol p (unpackInput_lofar,
    CDMT/Session_lofar_gpu.cu:144-175)."""
    _check(d_in, d_out, npol * nsamp * nchan, out_size(nbin, nfft, npol, nchan))
    call("sps_unpack_lofar", _dptr(d_in), int(nsamp), int(noverlap), int(nbin), int(nfft), int(npol), int(nchan),
         _dptr(d_out), _stream())
