"""Seeded synthetic filterbanks with injected dispersed pulses (generated on
the device with torch), for benchmarks and end-to-end tests.

A pulse of dispersion sweep D samples arrives at channel c (frequency f_c,
channel lower edge) delayed by D * (f_c^-2 - f_max^-2) / (f_min^-2 - f_max^-2)
samples relative to the top of the band, with a Gaussian time profile.
"""
import math

import numpy as np


def channel_freqs(f_min, f_max, nchans):
    df = (f_max - f_min) / nchans
    return f_min + df * np.arange(nchans, dtype=np.float64)


def pulse_list(seed, npulses, nsamps, max_sweep, width_range, snr_range, min_sweep=0):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(npulses):
        sweep = int(rng.integers(min_sweep, max_sweep + 1))
        width = float(rng.uniform(*width_range))
        snr = float(rng.uniform(*snr_range))
        t0 = int(rng.integers(0, max(1, nsamps - sweep - 4 * int(width) - 8)))
        out.append((t0, sweep, width, snr))
    return out


def inject(x, f_min, f_max, pulses, sigma, device="cuda"):
    """Adds pulses in place to x (nchans x nsamps float32 torch tensor). The
    per-channel per-sample peak amplitude is snr * sigma / sqrt(nchans * width)
    so a matched detection has roughly the stated S/N."""
    import torch

    nchans, nsamps = x.shape
    f = channel_freqs(f_min, f_max, nchans)
    k = (f ** -2 - f_max ** -2) / (f_min ** -2 - f_max ** -2)
    for (t0, sweep, width, snr) in pulses:
        delay = torch.from_numpy(np.round(k * sweep).astype(np.int64)).to(device)
        half = int(math.ceil(3 * width))
        offs = torch.arange(-half, half + 1, device=device)
        prof = torch.exp(-0.5 * (offs.float() / width) ** 2)
        amp = snr * sigma / math.sqrt(nchans * max(width, 1.0))
        t = t0 + delay[:, None] + offs[None, :]
        valid = (t >= 0) & (t < nsamps)
        rows = torch.arange(nchans, device=device)[:, None].expand_as(t)
        vals = (amp * prof)[None, :].expand_as(t)
        x.index_put_((rows[valid], t[valid]), vals[valid], accumulate=True)
    return x


def u8_filterbank(nchans, nsamps, f_min, f_max, seed, npulses=20, max_sweep=4096, width_range=(1.0, 256.0),
                  snr_range=(10.0, 30.0), mean=128.0, sigma=16.0, device="cuda", chunk_rows=512):
    """uint8 filterbank: clip(round(N(mean, sigma) + pulses), 0, 255)."""
    import torch

    g = torch.Generator(device=device)
    g.manual_seed(seed)
    out = torch.empty((nchans, nsamps), dtype=torch.uint8, device=device)
    pulses = pulse_list(seed, npulses, nsamps, max_sweep, width_range, snr_range)
    for r0 in range(0, nchans, chunk_rows):
        r1 = min(nchans, r0 + chunk_rows)
        x = torch.randn((r1 - r0, nsamps), generator=g, device=device) * sigma + mean
        out[r0:r1] = x.round_().clamp_(0, 255).to(torch.uint8)
        del x
    # pulses are added on a float copy of the affected region only
    xf = out.float()
    inject(xf, f_min, f_max, pulses, sigma, device)
    out.copy_(xf.round_().clamp_(0, 255).to(torch.uint8))
    del xf
    return out, pulses


def f32_filterbank(nchans, nsamps, f_min, f_max, seed, npulses=3, max_sweep=512, width_range=(1.0, 16.0),
                   snr_range=(10.0, 30.0), min_sweep=0, rfi_frac=0.0, rfi_sigma=5.0, device="cuda"):
    """float32 N(0,1) filterbank with pulses; a fraction of channels can be
    replaced by N(0, rfi_sigma^2) noise."""
    import torch

    g = torch.Generator(device=device)
    g.manual_seed(seed)
    x = torch.randn((nchans, nsamps), generator=g, device=device)
    pulses = pulse_list(seed, npulses, nsamps, max_sweep, width_range, snr_range, min_sweep)
    inject(x, f_min, f_max, pulses, 1.0, device)
    if rfi_frac > 0:
        rng = np.random.default_rng(seed + 1000)
        bad = rng.choice(nchans, int(round(rfi_frac * nchans)), replace=False)
        for c in bad:
            x[int(c)] = torch.randn(nsamps, generator=g, device=device) * rfi_sigma
    return x, pulses
