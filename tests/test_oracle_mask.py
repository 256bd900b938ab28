"""The premise of the search's counted mask transform (k_mask_counts),
checked on the oracle's FDMT (running-mean init) of a row-constant 0/1 keep
mask: every value is an exact integer, each row is non-decreasing in t (a
kept channel's contribution switches on at its accumulated shift and stays),
and every row reaches the kept-channel count once t passes its longest
track; with all channels kept it equals the FDMT of ones."""
import numpy as np

from oracle import orc


def test_mask_transform_counts_kept_channels():
    rng = np.random.default_rng(11)
    for nch, ns, dtm in ((64, 256, 63), (128, 600, 100), (256, 512, 255)):
        plan = orc.Plan(1200.0, 1600.0, nch, ns, 1.0, dtm)
        for frac in (1.0, 0.7):
            keep = (rng.random(nch) < frac).astype(np.float32)
            mask = np.repeat(keep[:, None], ns, axis=1)
            out = plan.execute(mask)
            assert np.array_equal(out, np.rint(out))
            assert np.all(np.diff(out, axis=1) >= 0)
            assert np.all(out[:, -1] == keep.sum())
            if frac == 1.0:
                np.testing.assert_array_equal(out, plan.execute(np.ones((nch, ns), np.float32)))
