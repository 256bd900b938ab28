"""Pins the oracle's 8-bit GUPPI / LOFAR raw-voltage unpack.

The reference's own example (unpack_guppi/kernel.cu:192-216: nbin 6, npol 4,
nchan 3, Noverlap 1, nfft 2, input 1 + i) is checked against the blocks read
straight off the input array's layout (numpy reshapes and slices, not the
index arithmetic of the restatement); the two layouts are checked against
each other through a transpose of the input; the overlap-save structure and
the zero fill are checked as properties.
"""
import numpy as np
import pytest

from oracle import orc

EX = dict(nbin=6, npol=4, nchan=3, noverlap=1, nfft=2)


def example_input():
    nsamp = EX["nfft"] * (EX["nbin"] - 2 * EX["noverlap"])
    n = nsamp * EX["nchan"] * EX["npol"]
    return nsamp, (1 + np.arange(n)).astype(np.int8)


def blocks_from_samples(z, nbin, nfft, noverlap):
    """z[pol][chan][samp] complex -> [pol][fft][chan][bin], zero outside the
    chunk: block f starts at sample f * (nbin - 2 * noverlap) - noverlap."""
    npol2, nchan, nsamp = z.shape
    pad = np.zeros((npol2, nchan, nbin + noverlap + nsamp + nbin + nfft * (nbin - 2 * noverlap)), np.complex64)
    pad[:, :, nbin + noverlap:nbin + noverlap + nsamp] = z  # sample s at column nbin + noverlap + s
    out = np.zeros((npol2, nfft, nchan, nbin), np.complex64)
    step = nbin - 2 * noverlap
    for f in range(nfft):
        c0 = nbin + f * step  # column of sample f * step - noverlap
        out[:, f] = pad[:, :, c0:c0 + nbin]
    return out


def test_guppi_reference_example():
    nsamp, raw = example_input()
    got = orc.unpack("guppi", raw, nsamp, **EX)
    x = raw.reshape(EX["nchan"], nsamp, EX["npol"]).astype(np.float32)  # [chan][samp][re0 im0 re1 im1]
    z = np.stack([x[:, :, 2 * p] + 1j * x[:, :, 2 * p + 1] for p in range(EX["npol"] // 2)])
    np.testing.assert_array_equal(got, blocks_from_samples(z, EX["nbin"], EX["nfft"], EX["noverlap"]))
    # block 0 starts one sample before the chunk, block 1 ends one after it
    assert (got[:, 0, :, 0] == 0).all() and (got[:, 1, :, -1] == 0).all()
    assert np.count_nonzero(got) == got.size - 2 * 2 * EX["nchan"]
    assert got[0, 0, 0, 1] == 1 + 2j and got[1, 0, 0, 1] == 3 + 4j  # the first sample of channel 0


def test_lofar_reference_example():
    nsamp, raw = example_input()
    got = orc.unpack("lofar", raw, nsamp, **EX)
    x = raw.reshape(EX["npol"], nsamp, EX["nchan"]).astype(np.float32)  # [component][samp][chan]
    z = np.stack([(x[2 * p] + 1j * x[2 * p + 1]).T for p in range(EX["npol"] // 2)])
    np.testing.assert_array_equal(got, blocks_from_samples(z, EX["nbin"], EX["nfft"], EX["noverlap"]))
    assert got[0, 0, 0, 1] == 1 + 25j  # re of sample 0 channel 0 is input 0, im is input nsamp * nchan


@pytest.mark.parametrize("nbin,nfft,noverlap,npol,nchan,nsamp",
                         [(16, 3, 2, 4, 5, 36), (9, 4, 0, 2, 7, 30), (12, 2, 3, 4, 2, 100), (8, 5, 1, 4, 3, 10)])
def test_layouts_agree_and_blocks_overlap(nbin, nfft, noverlap, npol, nchan, nsamp):
    rng = np.random.default_rng(nbin * 100 + nfft)
    g = rng.integers(-128, 128, (nchan, nsamp, npol), dtype=np.int8)
    lof = np.ascontiguousarray(g.transpose(2, 1, 0))  # [component][samp][chan]
    a = orc.unpack("guppi", g, nsamp, noverlap, nbin, nfft, npol, nchan)
    b = orc.unpack("lofar", lof, nsamp, noverlap, nbin, nfft, npol, nchan)
    np.testing.assert_array_equal(a, b)
    # signed promotion: the extremes survive
    z = np.stack([g[:, :, 2 * p] + 1j * g[:, :, 2 * p + 1] for p in range(npol // 2)]).astype(np.complex64)
    np.testing.assert_array_equal(a, blocks_from_samples(z, nbin, nfft, noverlap))
    # overlap-save: the last 2*noverlap bins of a block are the first of the next
    if noverlap:
        np.testing.assert_array_equal(a[:, :-1, :, nbin - 2 * noverlap:], a[:, 1:, :, :2 * noverlap])
    # every sample inside the chunk and covered by some block appears
    step = nbin - 2 * noverlap
    for s in range(min(nsamp, nfft * step + noverlap)):
        f = max(0, (s + noverlap) // step - (1 if (s + noverlap) // step >= nfft else 0))
        f = min(f, nfft - 1)
        b_ = s + noverlap - f * step
        if 0 <= b_ < nbin:
            assert a[0, f, 0, b_] == z[0, 0, s]
