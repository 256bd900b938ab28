"""Merge schedules: every split of the iterations into fused passes of depth
1..3 gives bit-identical output (golden fixtures and the oracle), including
plans with copied sub-bands (odd counts) and the legacy CFdmt plan."""
import itertools

import numpy as np
import pytest

from golden_inputs import make_input, sha
from oracle import orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def compositions(n, parts=(1, 2, 3)):
    if n == 0:
        yield []
        return
    for p in parts:
        if p <= n:
            for rest in compositions(n - p, parts):
                yield [p] + rest


def run(torch, f, x):
    rows, cols = f.output_shape()
    d_in = torch.from_numpy(np.ascontiguousarray(x)).cuda()
    d_out = torch.full((rows, cols), float("nan"), dtype=torch.float32, device="cuda")
    f.execute(d_in, d_in.numel(), d_out, d_out.numel())
    torch.cuda.synchronize()
    return d_out.cpu().numpy()


def test_all_schedules_small_odd_plan(torch_cuda):
    from cpp_projects_amd.fdmt import FDMTGPU

    nch, ns, dtm = 37, 300, 36
    p = orc.Plan(400.0, 800.0, nch, ns, 1.0, dtm)
    x = np.random.default_rng(0).standard_normal((nch, ns)).astype(np.float32)
    ref = p.execute(x)
    f = FDMTGPU(400.0, 800.0, nch, ns, 1.0, dtm)
    scheds = list(compositions(p.niters))
    assert len(scheds) > 10
    for s in scheds:
        f.set_schedule(s)
        assert f.get_schedule() == s
        np.testing.assert_array_equal(run(torch_cuda, f, x).view(np.uint32), ref.view(np.uint32), err_msg=str(s))


def test_schedules_on_golden(torch_cuda, golden):
    from cpp_projects_amd.fdmt import FDMTGPU

    man, arr = golden
    for name in ("c1000odd", "c1024", "c3000", "c256step2"):
        e = man["fdmt"][name]
        x = make_input(e["input"], e["nchans"], e["nsamps"], e["seed"], arr)
        f = FDMTGPU(e["fmin"], e["fmax"], e["nchans"], e["nsamps"], 1.0, e["dt_max"], e["dt_step"], e["dt_min"])
        n = e["niters"]
        cands = [[1] * n, [3] * (n // 3) + [n % 3] * (n % 3 > 0), [2] * (n // 2) + [1] * (n % 2),
                 [1, 3] + [2] * ((n - 4) // 2) + [1] * ((n - 4) % 2)]
        for s in cands:
            s = [d for d in s if d > 0]
            assert sum(s) == n
            f.set_schedule(s)
            print("case", name, s, flush=True)
            assert sha(run(torch_cuda, f, x)) == e["out_sha256"], (name, s)


def test_schedules_cfdmt(torch_cuda, golden):
    from cpp_projects_amd.fdmt import CFdmtGpu

    torch = torch_cuda
    man, arr = golden
    e = man["cfdmt"]["lc1000"]
    x = make_input(e["input"], e["nchan"], e["cols"], e["seed"], arr)
    f = CFdmtGpu(e["fmin"], e["fmax"], e["nchan"], e["cols"], e["imaxdt"])
    n = f.get_niters()
    d_in = torch.from_numpy(x).cuda()
    d_out = torch.empty((e["imaxdt"], e["cols"]), dtype=torch.float32, device="cuda")
    for s in ([3] * (n // 3) + [1] * (n % 3), [2] * (n // 2) + [1] * (n % 2), [1] * n):
        f.set_schedule(s)
        f.process_image(d_in, d_out, False)
        torch.cuda.synchronize()
        assert sha(d_out.cpu().numpy()) == e["out_sha256"], s


def test_bad_schedules_rejected(torch_cuda):
    from cpp_projects_amd._lib import InvalidArgument
    from cpp_projects_amd.fdmt import FDMTGPU

    f = FDMTGPU(1200.0, 1600.0, 64, 256, 1.0, 63)
    with pytest.raises(InvalidArgument):
        f.set_schedule([4, 2])
    with pytest.raises(InvalidArgument):
        f.set_schedule([3, 2])
    assert sum(f.get_schedule()) == f.get_niters()


def test_search_schedules_same_candidates(torch_cuda):
    from cpp_projects_amd.search import SingleBeamSearch, cands_to_numpy

    torch = torch_cuda
    rng = np.random.default_rng(9)
    nch, ns = 256, 8192
    x = np.clip(np.round(rng.normal(128, 16, (nch, ns))), 0, 255).astype(np.uint8)
    x[:, 3000:3004] = 200
    s = SingleBeamSearch(1200.0, 1600.0, nch, ns, nch - 1, in_dtype="u8", box_q=6, threshold=5.0,
                         cand_capacity=1 << 18)
    d_in = torch.from_numpy(x).cuda()
    dmt = torch.empty((s.ndt, ns), device="cuda")
    c = torch.empty((1 << 18, 4), dtype=torch.int32, device="cuda")
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    outs = []
    for sched in ([1] * 8, [3, 3, 2], [2, 2, 2, 2], [1, 3, 3, 1]):
        s.set_schedule(sched)
        s.run(d_in, dmt, c, cnt)
        n = s.sort_candidates(c, cnt)
        outs.append((sha(dmt.cpu().numpy()), cands_to_numpy(c, n).tobytes()))
    assert all(o == outs[0] for o in outs)


def test_staged_all_schedules_small_odd_plan(torch_cuda):
    from cpp_projects_amd.fdmt import FDMTGPU

    nch, ns, dtm = 37, 300, 36
    p = orc.Plan(400.0, 800.0, nch, ns, 1.0, dtm)
    x = np.random.default_rng(1).standard_normal((nch, ns)).astype(np.float32)
    ref = p.execute(x)
    f = FDMTGPU(400.0, 800.0, nch, ns, 1.0, dtm)
    f.set_staged(True)
    scheds = list(compositions(p.niters, (1, 2, 3, 4)))
    assert any(4 in s for s in scheds)
    for s in scheds:
        f.set_schedule(s)
        assert f.get_schedule() == s
        np.testing.assert_array_equal(run(torch_cuda, f, x).view(np.uint32), ref.view(np.uint32), err_msg=str(s))


def test_staged_on_golden(torch_cuda, golden):
    from cpp_projects_amd.fdmt import FDMTGPU

    man, arr = golden
    for name in ("c1000odd", "c1024", "c3000", "c256step2", "c512file"):
        e = man["fdmt"][name]
        x = make_input(e["input"], e["nchans"], e["nsamps"], e["seed"], arr)
        f = FDMTGPU(e["fmin"], e["fmax"], e["nchans"], e["nsamps"], 1.0, e["dt_max"], e["dt_step"], e["dt_min"])
        f.set_staged(True)
        n = e["niters"]
        cands = [[4] * (n // 4) + [n % 4] * (n % 4 > 0), [1] * n, [3] * (n // 3) + [n % 3] * (n % 3 > 0),
                 [2, 4] + [3] * ((n - 6) // 3) + [(n - 6) % 3] * ((n - 6) % 3 > 0)]
        for s in cands:
            s = [d for d in s if d > 0]
            assert sum(s) == n
            f.set_schedule(s)
            print("case staged", name, s, flush=True)
            assert sha(run(torch_cuda, f, x)) == e["out_sha256"], (name, s)


def test_staged_long_rows_vs_oracle(torch_cuda):
    """Many time tiles, a length that is not a multiple of 4 or of the tile,
    and a wide delay range (groups whose leaves span large shifts)."""
    from cpp_projects_amd.fdmt import FDMTGPU

    for nch, ns, dtm in ((512, 5001, 511), (256, 3000, 1000), (1024, 2050, 1023)):
        p = orc.Plan(1200.0, 1600.0, nch, ns, 1.0, dtm)
        x = np.random.default_rng(nch).standard_normal((nch, ns)).astype(np.float32)
        ref = p.execute(x)
        f = FDMTGPU(1200.0, 1600.0, nch, ns, 1.0, dtm)
        f.set_staged(True)
        n = p.niters
        for s in ([4] * (n // 4) + [n % 4] * (n % 4 > 0), [2] * (n // 2) + [1] * (n % 2)):
            s = [d for d in s if d > 0]
            f.set_schedule(s)
            print("case staged long", nch, ns, dtm, s, flush=True)
            np.testing.assert_array_equal(run(torch_cuda, f, x).view(np.uint32), ref.view(np.uint32),
                                          err_msg=f"{nch} {ns} {s}")


def test_staged_cfdmt_and_toggle(torch_cuda, golden):
    from cpp_projects_amd._lib import InvalidArgument
    from cpp_projects_amd.fdmt import CFdmtGpu

    torch = torch_cuda
    man, arr = golden
    e = man["cfdmt"]["lc1000"]
    x = make_input(e["input"], e["nchan"], e["cols"], e["seed"], arr)
    f = CFdmtGpu(e["fmin"], e["fmax"], e["nchan"], e["cols"], e["imaxdt"])
    n = f.get_niters()
    d_in = torch.from_numpy(x).cuda()
    d_out = torch.empty((e["imaxdt"], e["cols"]), dtype=torch.float32, device="cuda")
    f.set_staged(True)
    f.set_schedule([4] * (n // 4) + [1] * (n % 4))
    f.process_image(d_in, d_out, False)
    torch.cuda.synchronize()
    assert sha(d_out.cpu().numpy()) == e["out_sha256"]
    # leaving staged mode drops a schedule with depth-4 passes
    f.set_staged(False)
    assert max(f.get_schedule()) <= 3 and sum(f.get_schedule()) == n
    with pytest.raises(InvalidArgument):
        f.set_schedule([4] * (n // 4) + [1] * (n % 4))
    f.process_image(d_in, d_out, False)
    torch.cuda.synchronize()
    assert sha(d_out.cpu().numpy()) == e["out_sha256"]


def test_search_staged_same_candidates(torch_cuda):
    from cpp_projects_amd.search import SingleBeamSearch, cands_to_numpy

    torch = torch_cuda
    rng = np.random.default_rng(19)
    nch, ns = 256, 8190
    x = np.clip(np.round(rng.normal(128, 16, (nch, ns))), 0, 255).astype(np.uint8)
    x[:, 3000:3004] = 200
    s = SingleBeamSearch(1200.0, 1600.0, nch, ns, nch - 1, in_dtype="u8", box_q=6, threshold=5.0,
                         cand_capacity=1 << 18)
    d_in = torch.from_numpy(x).cuda()
    dmt = torch.empty((s.ndt, ns), device="cuda")
    c = torch.empty((1 << 18, 4), dtype=torch.int32, device="cuda")
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    outs = []
    for staged, sched in ((False, [3, 3, 2]), (True, [4, 4]), (True, [1, 3, 4]), (True, [2, 2, 2, 2])):
        s.set_staged(staged)
        s.set_schedule(sched)
        s.run(d_in, dmt, c, cnt)
        n = s.sort_candidates(c, cnt)
        outs.append((sha(dmt.cpu().numpy()), cands_to_numpy(c, n).tobytes()))
    assert all(o == outs[0] for o in outs)


def test_fused_init_on_golden(torch_cuda, golden):
    """Staging mode 2: the first pass builds the mean-init rows itself."""
    from cpp_projects_amd.fdmt import FDMTGPU

    man, arr = golden
    for name in ("c1000odd", "c1024", "c3000", "c256step2", "c512file", "c64"):
        e = man["fdmt"][name]
        x = make_input(e["input"], e["nchans"], e["nsamps"], e["seed"], arr)
        f = FDMTGPU(e["fmin"], e["fmax"], e["nchans"], e["nsamps"], 1.0, e["dt_max"], e["dt_step"], e["dt_min"])
        f.set_staged(2)
        n = e["niters"]
        for first in (1, 2, 3, 4):
            rest = n - first
            s = [first] + [3] * (rest // 3) + [rest % 3] * (rest % 3 > 0)
            s = [d for d in s if d > 0]
            f.set_schedule(s)
            print("case fused init", name, s, flush=True)
            assert sha(run(torch_cuda, f, x)) == e["out_sha256"], (name, s)


def test_fused_init_u8_and_restrict_fallback(torch_cuda):
    from cpp_projects_amd.fdmt import SPS_INIT_RESTRICT, FDMTGPU

    torch = torch_cuda
    rng = np.random.default_rng(23)
    nch, ns, dtm = 300, 4099, 299
    x = rng.integers(0, 256, (nch, ns), dtype=np.uint8)
    p = orc.Plan(1200.0, 1600.0, nch, ns, 1.0, dtm)
    ref = p.execute(x.astype(np.float32))
    f = FDMTGPU(1200.0, 1600.0, nch, ns, 1.0, dtm)
    f.set_staged(2)
    n = p.niters
    for s in ([3] * (n // 3) + [n % 3] * (n % 3 > 0), [4, 4] + [1] * (n - 8)):
        s = [d for d in s if d > 0]
        f.set_schedule(s)
        d_in = torch.from_numpy(x).cuda()
        d_out = torch.empty((dtm + 1, ns), dtype=torch.float32, device="cuda")
        f.execute(d_in, d_in.numel(), d_out, d_out.numel())
        torch.cuda.synchronize()
        np.testing.assert_array_equal(d_out.cpu().numpy().view(np.uint32), ref.view(np.uint32), err_msg=str(s))
    # RESTRICT init keeps the separate init kernel under mode 2
    xf = rng.standard_normal((nch, ns)).astype(np.float32)
    rr = orc.Plan(1200.0, 1600.0, nch, ns, 1.0, dtm).execute(xf, init_mode=1)
    g = FDMTGPU(1200.0, 1600.0, nch, ns, 1.0, dtm, init_mode=SPS_INIT_RESTRICT)
    g.set_staged(2)
    g.set_schedule([3] * (n // 3) + [n % 3] * (n % 3 > 0) if n % 3 else [3] * (n // 3))
    np.testing.assert_array_equal(run(torch, g, xf).view(np.uint32), rr.view(np.uint32))


def test_search_fused_init_vs_oracle_chain(torch_cuda):
    from cpp_projects_amd.search import SingleBeamSearch, cands_to_numpy

    torch = torch_cuda
    rng = np.random.default_rng(29)
    nch, ns, dtm = 256, 6000, 255
    x = np.clip(np.round(rng.normal(128, 16, (nch, ns))), 0, 255).astype(np.uint8)
    x[17] = np.clip(np.round(rng.normal(128, 70, ns)), 0, 255).astype(np.uint8)  # RFI row
    x[:, 2000:2003] = 210
    s = SingleBeamSearch(1200.0, 1600.0, nch, ns, dtm, in_dtype="u8", box_q=5, threshold=5.0, cand_capacity=1 << 17)
    s.set_staged(2)
    s.set_schedule([4, 4])
    d_in = torch.from_numpy(x).cuda()
    dmt = torch.empty((s.ndt, ns), dtype=torch.float32, device="cuda")
    c = torch.empty((1 << 17, 4), dtype=torch.int32, device="cuda")
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    s.run(d_in, dmt, c, cnt)
    k = s.sort_candidates(c, cnt)
    xn, mask, (_, _, keep, _) = orc.normalise(x, 1)
    assert keep[17] == 0
    p = orc.Plan(1200.0, 1600.0, nch, ns, 1.0, dtm)
    snr = orc.fdmt_normalization(p.execute(xn), p.execute(mask))
    np.testing.assert_array_equal(dmt.cpu().numpy().view(np.uint32), snr.view(np.uint32))
    ref = orc.boxcar_pow2(snr, 5.0, 5)
    assert k == len(ref) and k > 0
    np.testing.assert_array_equal(cands_to_numpy(c, k).view(np.int32), ref.view(np.int32))


@pytest.mark.parametrize("nch,ns,dtm", [(512, 8192, 511), (1000, 6000, 999), (64, 4096, 63), (300, 5003, 299),
                                        (257, 4097, 256)])
def test_fused_init_streaming_f32(torch_cuda, nch, ns, dtm):
    """The streaming fused init (modes 2 and 4, init dts 0..K-1 with K = 2 or 3)
    over several tile counts per block and first-pass depths, bit-exact
    against the oracle: whole-chunk loads when nsamps is a multiple of 4,
    sample loads otherwise (rows then start unaligned) and for a waterfall
    that is not 16-byte aligned."""
    from cpp_projects_amd.fdmt import FDMTGPU

    torch = torch_cuda
    rng = np.random.default_rng(nch + ns)
    x = rng.standard_normal((nch, ns)).astype(np.float32)
    x[:, 100:103] += 4.0
    ref = orc.Plan(1200.0, 1600.0, nch, ns, 1.0, dtm).execute(x)
    f = FDMTGPU(1200.0, 1600.0, nch, ns, 1.0, dtm)
    n = f.get_niters()
    for mode in (2, 4):
        f.set_staged(mode)
        for tiles in (1, 5, 32):
            f.set_group_shape(6144, 16, tiles)
            for first in (1, 2, 3):
                rest = n - first
                sched = [first] + [3] * (rest // 3) + ([rest % 3] if rest % 3 else [])
                f.set_schedule(sched)
                got = run(torch, f, x)
                np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32),
                                              err_msg=f"mode {mode} tiles {tiles} schedule {sched}")
    # a view one float into a buffer: 4-byte aligned only
    buf = torch.zeros(nch * ns + 4, dtype=torch.float32, device="cuda")
    buf[1:1 + nch * ns] = torch.from_numpy(x.reshape(-1)).cuda()
    view = buf[1:1 + nch * ns]
    out = torch.empty((dtm + 1, ns), dtype=torch.float32, device="cuda")
    f.execute(view, view.numel(), out, out.numel())
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out.cpu().numpy().view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("in_dtype,ns", [("u8", 16384), ("f32", 16384), ("u8", 16381), ("f32", 16383)])
def test_search_streaming_fused_init_vs_oracle(torch_cuda, in_dtype, ns):
    """The search with the init fused into its first staged pass (mode 4):
    normalisation applied while the waterfall is read, bit-exact S/N image and
    candidates, also for lengths whose rows start unaligned; the init stage
    then takes no time of its own."""
    from cpp_projects_amd.search import SingleBeamSearch, cands_to_numpy

    torch = torch_cuda
    rng = np.random.default_rng(31)
    nch, dtm = 1024, 1023
    if in_dtype == "u8":
        x = np.clip(np.round(rng.normal(128, 16, (nch, ns))), 0, 255).astype(np.uint8)
        x[77] = np.clip(np.round(rng.normal(128, 90, ns)), 0, 255).astype(np.uint8)  # a row the cleaning drops
        x[:, 5000:5003] = 200
    else:
        x = rng.standard_normal((nch, ns)).astype(np.float32)
        x[77] *= 40.0
        x[:, 5000:5003] += 3.0
    s = SingleBeamSearch(1200.0, 1600.0, nch, ns, dtm, in_dtype=in_dtype, box_q=5, threshold=5.0,
                         cand_capacity=1 << 18)
    s.set_staged(4)
    s.set_group_shape(6144, 16, 8)
    s.set_schedule([3, 3, 3, 1])
    d_in = torch.from_numpy(x).cuda()
    dmt = torch.empty((s.ndt, ns), dtype=torch.float32, device="cuda")
    c = torch.empty((1 << 18, 4), dtype=torch.int32, device="cuda")
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    s.enable_timing(True)
    s.run(d_in, dmt, c, cnt)
    st = s.stage_times_ms()
    s.enable_timing(False)
    assert st[2] < 0.02, f"init stage {st[2]} ms: the fused first pass did not run"
    ct, k = s.candidates(dmt, c, cnt)
    xn, mask, (_, _, keep, _) = orc.normalise(x, 1)
    assert keep[77] == 0
    p = orc.Plan(1200.0, 1600.0, nch, ns, 1.0, dtm)
    snr = orc.fdmt_normalization(p.execute(xn), p.execute(mask))
    np.testing.assert_array_equal(dmt.cpu().numpy().view(np.uint32), snr.view(np.uint32))
    ref = orc.boxcar_pow2(snr, 5.0, 5)
    assert k == len(ref) and k > 0
    np.testing.assert_array_equal(cands_to_numpy(ct, k).view(np.int32), ref.view(np.int32))


def test_search_fused_init_flat_u8(torch_cuda):
    """A flat 8-bit waterfall: the global sigma is 0, so no channel passes the
    fused init's range test and the pass takes its checked branch (every
    sample through the IEEE division); the S/N image holds the oracle's NaNs
    and zeros, and there are no candidates."""
    from cpp_projects_amd.search import SingleBeamSearch

    torch = torch_cuda
    nch, ns, dtm = 256, 8192, 255
    x = np.full((nch, ns), 100, np.uint8)
    s = SingleBeamSearch(1200.0, 1600.0, nch, ns, dtm, in_dtype="u8", box_q=4, threshold=5.0, cand_capacity=1 << 10)
    s.set_staged(4)
    s.set_schedule([3, 3, 2])
    d_in = torch.from_numpy(x).cuda()
    dmt = torch.empty((s.ndt, ns), dtype=torch.float32, device="cuda")
    c = torch.empty((1 << 10, 4), dtype=torch.int32, device="cuda")
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    s.enable_timing(True)
    s.run(d_in, dmt, c, cnt)
    st = s.stage_times_ms()
    s.enable_timing(False)
    assert st[2] < 0.02, f"init stage {st[2]} ms: the fused first pass did not run"
    k = s.sort_candidates(c, cnt)
    xn, mask, (_, _, keep, stv) = orc.normalise(x, 1)
    assert stv[1] == 0.0 and keep.all()
    p = orc.Plan(1200.0, 1600.0, nch, ns, 1.0, dtm)
    snr = orc.fdmt_normalization(p.execute(xn), p.execute(mask))
    got = dmt.cpu().numpy()
    assert np.isnan(snr).any()
    np.testing.assert_array_equal(np.isnan(got), np.isnan(snr))
    np.testing.assert_array_equal(got[~np.isnan(got)], snr[~np.isnan(snr)])
    assert k == 0 and len(orc.boxcar_pow2(snr, 5.0, 4)) == 0


@pytest.mark.parametrize("floats", [9216, 7168])
def test_staged_producer_consumer_vs_oracle(torch_cuda, floats):
    """Windows above 6144 floats run their fast tiles on the producer/consumer
    kernel (one loading wave, double-buffered LDS): many fast tiles per
    block, pass depths 1 to 4, with and without the S/N epilogue, bit-exact
    against the oracle."""
    from cpp_projects_amd.fdmt import FDMTGPU
    from cpp_projects_amd.search import SingleBeamSearch, cands_to_numpy

    torch = torch_cuda
    rng = np.random.default_rng(floats)
    nch, ns, dtm = 512, 32768, 511
    x = rng.standard_normal((nch, ns)).astype(np.float32)
    x[:, 7000:7004] += 3.0
    p = orc.Plan(1200.0, 1600.0, nch, ns, 1.0, dtm)
    ref = p.execute(x)
    f = FDMTGPU(1200.0, 1600.0, nch, ns, 1.0, dtm)
    for mode in (3, 4):
        f.set_staged(mode)
        for tiles in (1, 7, 64):
            f.set_group_shape(floats, 64, tiles)
            for sched in ([3, 3, 3], [4, 4, 1], [1, 4, 4], [2, 3, 4]):
                f.set_schedule(sched)
                print("case pc", floats, mode, tiles, sched, flush=True)
                got = run(torch, f, x)
                np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32),
                                              err_msg=f"mode {mode} tiles {tiles} schedule {sched}")
    xn, mask, _ = orc.normalise(x, 1)
    snr = orc.fdmt_normalization(p.execute(xn), p.execute(mask))
    cref = orc.boxcar_pow2(snr, 5.0, 5)
    s = SingleBeamSearch(1200.0, 1600.0, nch, ns, dtm, in_dtype="f32", box_q=5, threshold=5.0, cand_capacity=1 << 18)
    s.set_staged(4)
    s.set_group_shape(floats, 64, 16)
    d_in = torch.from_numpy(x).cuda()
    dmt = torch.empty((s.ndt, ns), dtype=torch.float32, device="cuda")
    c = torch.empty((1 << 18, 4), dtype=torch.int32, device="cuda")
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    for sched in ([3, 3, 3], [3, 4, 2], [2, 3, 4]):
        s.set_schedule(sched)
        print("case pc search", floats, sched, flush=True)
        cnt.zero_()
        s.run(d_in, dmt, c, cnt)
        k = s.sort_candidates(c, cnt)
        np.testing.assert_array_equal(dmt.cpu().numpy().view(np.uint32), snr.view(np.uint32), err_msg=str(sched))
        assert k == len(cref) and k > 0
        np.testing.assert_array_equal(cands_to_numpy(c, k).view(np.int32), cref.view(np.int32))


def test_staged_group_shapes(torch_cuda, golden):
    """Group windows from 2048 to 9216 floats (one LDS buffer above 6144), 1 to
    64 rows per group and 1 to 64 tiles per block give the same bits (plain
    staged and fused init, with and without a staged final pass)."""
    from cpp_projects_amd.fdmt import FDMTGPU

    man, arr = golden
    e = man["fdmt"]["c3000"]
    x = make_input(e["input"], e["nchans"], e["nsamps"], e["seed"], arr)
    f = FDMTGPU(e["fmin"], e["fmax"], e["nchans"], e["nsamps"], 1.0, e["dt_max"], e["dt_step"], e["dt_min"])
    n = e["niters"]
    for mode in (1, 2, 3, 4):
        f.set_staged(mode)
        for floats, rows, tiles in ((2048, 1, 1), (2048, 64, 3), (4096, 4, 8), (9216, 64, 64), (8192, 32, 5)):
            f.set_group_shape(floats, rows, tiles)
            for s in ([3] * (n // 3) + [n % 3] * (n % 3 > 0), [4] * (n // 4) + [n % 4] * (n % 4 > 0)):
                s = [d for d in s if d > 0]
                f.set_schedule(s)
                print("case shape", mode, floats, rows, tiles, s, flush=True)
                assert sha(run(torch_cuda, f, x)) == e["out_sha256"], (mode, floats, rows, tiles, s)
