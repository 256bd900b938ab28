"""The oracle's coherent front-end (oracle/orc.py coherent_*), checked on the
CPU against what the mathematics fixes: at DM 0 the chirp factor is the real
taper / nbin; the intensity pass only permutes and trims, so its total is
the total intensity of the kept samples; a tone lands in the row the index
maps of CDMT/Chunk_gpu.cu:476-515 predict (derived independently in the
test); and the dispersion shift moves a row by the rounded delay of its
frequency."""
import numpy as np

from oracle import orc


def test_chirp_factor_at_dm0_is_the_taper():
    nchan, len_sft, nbin, nfft, npol = 2, 4, 64, 1, 2
    ones = np.ones((npol // 2, nfft, nchan, nbin), np.complex64)
    f = orc.coherent_chirp(ones, 0.0, 1200.0, 1600.0, nchan, len_sft, nbin, nfft, npol)
    assert np.all(f.imag == 0)
    mbin = nbin // len_sft
    bw_chan = 400.0 / nchan / len_sft
    bf = -0.5 * bw_chan + (np.arange(mbin) + 0.5) * bw_chan / mbin
    taper = (1.0 / np.sqrt(1.0 + (bf / (0.47 * bw_chan)) ** 80) / nbin).astype(np.float32)
    for s in range(len_sft):
        cols = (s * mbin + np.arange(mbin) + nbin // 2) % nbin
        np.testing.assert_array_equal(f[0, 0, 1, cols].real, taper)


def test_detect_conserves_the_kept_intensity():
    rng = np.random.default_rng(4)
    npol, nchan, len_sft, nbin, nfft, nov = 4, 3, 8, 512, 2, 64
    buf = (rng.standard_normal((npol // 2, nfft, nchan, nbin)) +
           1j * rng.standard_normal((npol // 2, nfft, nchan, nbin))).astype(np.complex64)
    w = orc.coherent_detect(buf, 17.0, 1200.0, 1600.0, nchan, len_sft, nbin, nfft, npol, nov, 1e-6)
    mbin = nbin // len_sft
    nsub = nov // len_sft
    inten = (np.abs(buf.astype(np.complex128)) ** 2).sum(axis=0) / nbin / nbin  # [fft][chan][bin]
    # after the half-row rotation, segment s' holds block segment (s' - len_sft/2) mod len_sft
    kept = 0.0
    for s in range(len_sft):
        src = (s - len_sft // 2) % len_sft
        kept += inten[:, :, src * mbin + nsub:src * mbin + mbin - nsub].sum()
    np.testing.assert_allclose(w.astype(np.float64).sum(), kept, rtol=1e-5)


def test_tone_row_and_dispersion_shift():
    f_min, f_max, npol, nchan, len_sft, nbin, nfft, nov, tsamp = 1200.0, 1600.0, 2, 4, 8, 1024, 2, 64, 1e-6
    mbin = nbin // len_sft
    c, k = 1, 5 * mbin + 7
    x = np.zeros((npol // 2, nfft, nchan, nbin), np.complex64)
    x[:, :, c, :] = np.exp(2j * np.pi * k * np.arange(nbin) / nbin)
    w0 = orc.coherent_chain(x, 0.0, f_min, f_max, nchan, len_sft, nbin, nfft, npol, nov, tsamp)
    row = (nchan - 1 - c) * len_sft + len_sft - 1 - ((k // mbin + len_sft // 2) % len_sft)
    assert int(np.argmax(w0.sum(axis=1))) == row
    # an impulse in time after detection: the dispersion pass rotates row r by
    # round(4.148808e3 * dm * (fmin^-2 - f_r^-2) / (len_sft * tsamp))
    buf = np.zeros((npol // 2, nfft, nchan, nbin), np.complex64)
    buf[0, 0, :, :] = 1.0
    dm = 5.0
    a = orc.coherent_detect(buf, 0.0, f_min, f_max, nchan, len_sft, nbin, nfft, npol, nov, tsamp)
    b = orc.coherent_detect(buf, dm, f_min, f_max, nchan, len_sft, nbin, nfft, npol, nov, tsamp)
    nrows = nchan * len_sft
    foff = (f_max - f_min) / nrows
    tsw = float(np.float32(np.float32(len_sft) * np.float32(tsamp)))
    for r in (0, nrows // 2, nrows - 1):
        fr = (nrows - 1 - r) * foff + f_min
        sh = int(np.floor(4.148808e3 * dm * (1 / f_min ** 2 - 1 / fr ** 2) / tsw + 0.5))
        np.testing.assert_array_equal(b[r], np.roll(a[r], sh))


def test_mask_transform_is_causal_and_flat():
    """The premise of CoherentChunk.process's narrow mask transform
    (fdmt_normalization_rows): the CFdmt (imaxDt = L) of a row-constant 0/1
    keep mask equals, over its first W >= L columns, the CFdmt of those W
    columns alone, and from column W - 1 on it is constant along each row --
    bit for bit, at W = 4 L as the pipeline uses and at W = L."""
    rng = np.random.default_rng(5)
    for rows, L, cols in ((64, 8, 384), (256, 32, 1000), (128, 16, 300), (64, 8, 20)):
        keep = (rng.random(rows) > 0.2).astype(np.float32)
        mask = np.repeat(keep[:, None], cols, axis=1)
        full = orc.CFdmt(1200.0, 1600.0, rows, cols, L).process(mask)
        for W in (min(cols, 4 * L), L):
            narrow = orc.CFdmt(1200.0, 1600.0, rows, W, L).process(np.ascontiguousarray(mask[:, :W]))
            np.testing.assert_array_equal(full[:, :W], narrow)
            np.testing.assert_array_equal(full[:, W:], np.repeat(narrow[:, W - 1:W], cols - W, axis=1))
