"""Pins the CPU oracle's FDMT plan and execute against the reference.

Checks: golden fixtures written from the reference FDMTCPU / CFdmtCpu built
from source; the live reference build (oracle/_ref) where present; an FP64
brute-force shift-and-add along the plan tracks; the all-ones closed form;
linearity; and the reference's error paths.
"""
import numpy as np
import pytest

from golden_inputs import make_input, plan_tables, sha, tables_sha
from oracle import orc


def _plan(e):
    return orc.Plan(e["fmin"], e["fmax"], e["nchans"], e["nsamps"], 1.0, e["dt_max"], e["dt_step"], e["dt_min"])


def test_plan_tables_match_golden(golden):
    man, arr = golden
    for name, e in man["fdmt"].items():
        p = _plan(e)
        assert p.niters == e["niters"], name
        tabs = plan_tables(p)
        assert tables_sha(tabs) == e["plan_sha256"], name
        if f"{name}__grid0" in arr.files:
            for k, v in tabs.items():
                np.testing.assert_array_equal(v, arr[f"{name}__{k}"].reshape(v.shape), err_msg=f"{name} {k}")


@pytest.mark.parametrize("mode,key", [(0, "out_sha256"), (1, "restrict_out_sha256")])
def test_fdmt_execute_bitexact_golden(golden, mode, key):
    man, arr = golden
    for name, e in man["fdmt"].items():
        if key not in e:
            continue
        x = make_input(e["input"], e["nchans"], e["nsamps"], e["seed"], arr)
        out = _plan(e).execute(x, init_mode=mode)
        assert list(out.shape) == e["out_shape"], name
        if mode == 0 and f"{name}__out" in arr.files:
            np.testing.assert_array_equal(out.view(np.uint32), arr[f"{name}__out"].view(np.uint32), err_msg=name)
        assert sha(out) == e[key], name


def test_cfdmt_bitexact_golden(golden):
    man, arr = golden
    for name, e in man["cfdmt"].items():
        f = orc.CFdmt(e["fmin"], e["fmax"], e["nchan"], e["cols"], e["imaxdt"])
        assert f.niter == e["niter"]
        cfg = {}
        for i in range(f.niter + 1):
            cs, fr = f.config(i)
            cfg[f"cumsum{i}"] = cs
            cfg[f"freq{i}"] = fr
        assert tables_sha(cfg) == e["config_sha256"], name
        x = make_input(e["input"], e["nchan"], e["cols"], e["seed"], arr)
        out = f.process(x)
        ones = f.process(None, b_ones=True)
        assert sha(out) == e["out_sha256"], name
        assert sha(ones) == e["ones_sha256"], name
        if f"{name}__out" in arr.files:
            np.testing.assert_array_equal(out, arr[f"{name}__out"])


@pytest.mark.skipif(not orc.ref_available(), reason="reference build (oracle/_ref) not present")
def test_against_live_reference_random_configs():
    rng = np.random.default_rng(7)
    for _ in range(12):
        nch = int(rng.integers(2, 300))
        ns = int(rng.integers(64, 700))
        fmin = float(rng.uniform(100, 1500))
        fmax = fmin + float(rng.uniform(10, 800))
        dtm = int(rng.integers(1, min(ns - 1, 400)))
        step = int(rng.choice([1, 1, 2, 3]))
        dmin = int(rng.integers(0, max(1, dtm // 4)))
        try:
            r = orc.RefFDMT("pravir", fmin, fmax, nch, ns, 1.0, dtm, step, dmin)
        except RuntimeError:
            with pytest.raises(RuntimeError):
                orc.Plan(fmin, fmax, nch, ns, 1.0, dtm, step, dmin)
            continue
        p = orc.Plan(fmin, fmax, nch, ns, 1.0, dtm, step, dmin)
        assert tables_sha(plan_tables(p)) == tables_sha(plan_tables(r))
        x = rng.standard_normal((nch, ns)).astype(np.float32)
        np.testing.assert_array_equal(p.execute(x).view(np.uint32), r.execute(x).view(np.uint32))


def test_bruteforce_fp64_tracks():
    """fp32 FDMT vs FP64 sum over the plan's channel tracks: |err| <= 1e-5 of
    the cell's absolute magnitude scale (sum of |x| over the tracks is bounded
    by nchans * max|x|, so the bound used is 1e-5 * nchans * max|x| / nchans**0.5)."""
    for (nch, ns, dtm) in [(64, 512, 63), (37, 300, 36), (130, 400, 129)]:
        p = orc.Plan(1200.0, 1600.0, nch, ns, 1.0, dtm)
        x = np.random.default_rng(nch).standard_normal((nch, ns)).astype(np.float32)
        out = p.execute(x).astype(np.float64)
        bf = p.bruteforce(x)
        absbf = p.bruteforce(np.abs(x))
        err = np.abs(out - bf)
        assert (err <= 1e-5 * np.maximum(absbf, 1.0)).all(), err.max()


def test_ones_closed_form():
    """Every valid cell of the all-ones transform equals nchans exactly."""
    for (nch, ns, dtm, fmin, fmax) in [(64, 512, 63, 1200.0, 1600.0), (37, 300, 36, 400.0, 800.0), (100, 256, 99, 1200.0, 1600.0)]:
        p = orc.Plan(fmin, fmax, nch, ns, 1.0, dtm)
        o = p.execute(np.ones((nch, ns), np.float32))
        ext = p.track_extent()
        for r in range(o.shape[0]):
            assert (o[r, ext[r]:] == nch).all()
            if ext[r] > 0:
                assert o[r, ext[r] - 1] < nch


def test_impulse_lands_on_track():
    """A delta at (c, t0) contributes 1/(d+1) at t0 + shift + j, j in [0, d]."""
    nch, ns = 32, 256
    p = orc.Plan(1200.0, 1600.0, nch, ns, 1.0, 31)
    x = np.zeros((nch, ns), np.float32)
    x[5, 40] = 1.0
    o = p.execute(x)
    bf = p.bruteforce(x)
    np.testing.assert_allclose(o, bf, atol=1e-6)
    assert o.sum() > 0


def test_linearity():
    p = orc.Plan(1200.0, 1600.0, 48, 256, 1.0, 47)
    rng = np.random.default_rng(3)
    a = rng.integers(-8, 8, (48, 256)).astype(np.float32)
    b = rng.integers(-8, 8, (48, 256)).astype(np.float32)
    # integer inputs whose means have power-of-two denominators are exact;
    # compare within fp32 rounding
    np.testing.assert_allclose(p.execute(a + b), p.execute(a) + p.execute(b), atol=1e-4)


def test_error_paths():
    # offset >= nsamps -> "Offset is greater than input size" (fdmt_base.cpp:254-256)
    with pytest.raises(RuntimeError, match="Offset"):
        orc.Plan(1200.0, 1600.0, 64, 16, 1.0, 63)
    # legacy restrict init needs unit dt steps
    p = orc.Plan(1200.0, 1600.0, 16, 64, 1.0, 15, 2, 0)
    with pytest.raises(RuntimeError):
        p.execute(np.zeros((16, 64), np.float32), init_mode=1)
