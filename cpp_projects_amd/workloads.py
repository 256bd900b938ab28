"""The search workloads this package is measured and checked on, with their
seeded synthetic inputs (generated on the device).

Every workload covers 1200-1600 MHz with a channel's frequency at its lower
edge and row 0 the lowest frequency.

* ``PLAIN_CPU``: 512 ch x 32768 float32 N(0,1) + 3 pulses (sweep 50-450,
  widths 1-16), dt_max 511, boxcar widths 1..32 (q = 5), threshold 6,
  clustering bins {2, 1, 2}.
* ``SINGLE_BEAM``: 4096 ch x 262144 uint8, clip(round(N(128, 16))) + 20
  pulses (sweep up to 4096, widths 1-256), dt_max 4095, widths 1..256 (q = 8),
  threshold 6.
* ``BATCH``: 64 float32 beams, channel counts drawn from {1024, 2048, 4096},
  lengths uniform in 2^14..2^18, 0-5 pulses each, dt_max = nch - 1, widths
  1..256, threshold 6.
* ``LARGE``: 8192 ch x 2^20 float32, 10 % of the channels N(0, 25)
  interference, 50 pulses (sweep up to 8192, widths 1-512), dt_max 8191,
  widths 1..512 (q = 9), threshold 6.
"""
import numpy as np

from . import synth

F_MIN, F_MAX = 1200.0, 1600.0

PLAIN_CPU = dict(nchans=512, nsamps=1 << 15, dt_max=511, box_q=5, threshold=6.0, npulses=3, seed=1,
                 min_sweep=50, max_sweep=450, width_range=(1.0, 16.0), bins=(2, 1, 2))
SINGLE_BEAM = dict(nchans=4096, nsamps=1 << 18, dt_max=4095, box_q=8, threshold=6.0, npulses=20, seed=2)
BATCH = dict(nbeams=64, nch_choices=(1024, 2048, 4096), n_lo=1 << 14, n_hi=1 << 18, seed=3, box_q=8, threshold=6.0)
# the coherent front-end on a GUPPI-like chunk: 64 coarse channels of 6.25 MHz
# (1200-1600 MHz), two polarisations, 32 sub-channels each, 2^16-bin FFT
# blocks with 2^13 bins of overlap, 16 blocks per chunk, 8 coherent DMs
COHERENT = dict(f_min=1200.0, f_max=1600.0, npol=4, nchan=64, len_sft=32, nbin=1 << 16, nfft=16, noverlap=1 << 13,
                tsamp=1.6e-7, dms=(0.0, 25.0, 50.0, 75.0, 100.0, 125.0, 150.0, 175.0), seed=9,
                sigma_bound=6.0)
LARGE = dict(nchans=8192, nsamps=1 << 20, dt_max=8191, box_q=9, threshold=6.0, npulses=50, seed=5, rfi_frac=0.1,
             rfi_sigma=5.0)

# the staged-pass configuration the autotuner picks for SINGLE_BEAM on MI355X
SINGLE_BEAM_TUNED = dict(staged=4, shape=(9216, 64, 16), schedule=[3, 3, 3, 3])
# and for the large share
LARGE_TUNED = dict(staged=4, shape=(9216, 32, 16), schedule=[2, 2, 3, 3, 3])
# and for the batched beams of each channel count (tuned on the longest beam)
BATCH_TUNED = {1024: dict(staged=4, shape=(8192, 32, 16), schedule=[2, 2, 3, 3]),
               2048: dict(staged=4, shape=(8192, 64, 16), schedule=[2, 3, 3, 3]),
               4096: dict(staged=4, shape=(8192, 32, 16), schedule=[3, 3, 3, 3])}


def plain_cpu_input(device="cuda"):
    c = PLAIN_CPU
    return synth.f32_filterbank(c["nchans"], c["nsamps"], F_MIN, F_MAX, c["seed"], npulses=c["npulses"],
                                max_sweep=c["max_sweep"], min_sweep=c["min_sweep"], width_range=c["width_range"],
                                device=device)


def single_beam_input(device="cuda"):
    c = SINGLE_BEAM
    return synth.u8_filterbank(c["nchans"], c["nsamps"], F_MIN, F_MAX, c["seed"], npulses=c["npulses"],
                               device=device)


def batch_specs(nbeams=None):
    """Beam b: channel count from {1024, 2048, 4096}, length uniform in
    2^14..2^18 samples, 0-5 pulses, seeded 3 + b; dt_max = nch - 1."""
    rng = np.random.default_rng(BATCH["seed"])
    specs = []
    for b in range(BATCH["nbeams"] if nbeams is None else nbeams):
        nch = int(rng.choice(BATCH["nch_choices"]))
        ns = int(rng.integers(BATCH["n_lo"], BATCH["n_hi"] + 1))
        npulses = int(rng.integers(0, 6))
        specs.append(dict(nch=nch, ns=ns, npulses=npulses, seed=BATCH["seed"] + b))
    return specs


def batch_beam_input(spec, device="cuda"):
    x, _ = synth.f32_filterbank(spec["nch"], spec["ns"], F_MIN, F_MAX, spec["seed"], npulses=spec["npulses"],
                                max_sweep=spec["nch"] - 1, width_range=(1.0, 16.0), device=device)
    return x


def large_input(device="cuda"):
    c = LARGE
    return synth.f32_filterbank(c["nchans"], c["nsamps"], F_MIN, F_MAX, c["seed"], npulses=c["npulses"],
                                max_sweep=c["dt_max"], width_range=(1.0, 512.0), rfi_frac=c["rfi_frac"],
                                rfi_sigma=c["rfi_sigma"], device=device)
