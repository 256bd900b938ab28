"""GPU parity of normalisation, S/N normalisation, boxcar gathers, clustering,
the detector and the fused single-beam search through the C ABI."""
import csv
import os

import numpy as np
import pytest

from oracle import orc

pytestmark = pytest.mark.gpu
GD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def golden_image():
    img = np.full((3, 16), 2.0, np.float32)
    img[0, 4] = 140.0
    img[0, 11] = 120.0
    img[2, 4] = 100.0
    return img


def thr_tensor(torch, v):
    return torch.tensor([v], dtype=torch.float32, device="cuda")


def tuples(c):
    return [(int(a["mt"]), int(a["mdt"]), int(a["mwidth"]), float(a["msnr"])) for a in c]


def test_gather_candidates_golden(torch_cuda):
    from cpp_projects_amd.search import cands_to_numpy, clusterization

    torch = torch_cuda
    img = torch.from_numpy(golden_image()).cuda()
    thr = thr_tensor(torch, 4.0)
    ok, c2, n2 = clusterization.gather_candidates_in_heap_with_plan(img, 3, 16, thr, 2)
    got2 = tuples(cands_to_numpy(c2, n2))
    with open(os.path.join(GD, "Candidates2.csv")) as f:
        exp2 = [(int(r["mt"]), int(r["mdt"]), int(r["mwidth"])) for r in csv.DictReader(f)]
    assert ok and [g[:3] for g in got2] == exp2
    assert got2 == tuples(orc.boxcar_all(golden_image(), 4.0, 2))
    c1, n1 = clusterization.gather_candidates_in_dynamicalArray(img, 3, 16, thr, 1)
    got1 = tuples(cands_to_numpy(c1, n1))
    with open(os.path.join(GD, "Candidates1.csv")) as f:
        exp1 = sorted((int(r["mt"]), int(r["mdt"]), int(r["mwidth"])) for r in csv.DictReader(f))
    assert sorted(g[:3] for g in got1) == exp1
    assert got1 == tuples(orc.boxcar_pow2(golden_image(), 4.0, 1))


def test_boxcar_random_vs_oracle(torch_cuda):
    from cpp_projects_amd.search import cands_to_numpy, clusterization

    torch = torch_cuda
    rng = np.random.default_rng(4)
    for (R, N, q, thr) in [(7, 3000, 4, 3.0), (40, 1500, 6, 3.5), (3, 5000, 8, 2.5)]:
        img = rng.standard_normal((R, N)).astype(np.float32)
        img[R // 2, N // 3: N // 3 + 40] += 1.5
        d = torch.from_numpy(img).cuda()
        c, n = clusterization.gather_candidates_in_dynamicalArray(d, R, N, thr_tensor(torch, thr), q)
        ref = orc.boxcar_pow2(img, thr, q)
        assert n == len(ref)
        got = cands_to_numpy(c, n)
        np.testing.assert_array_equal(got.view(np.int32), ref.view(np.int32))
        ok, c, n = clusterization.gather_candidates_in_heap_with_plan(d, R, N, thr_tensor(torch, thr), 9)
        ref = orc.boxcar_all(img, thr, 9)
        assert n == len(ref)
        np.testing.assert_array_equal(cands_to_numpy(c, n).view(np.int32), ref.view(np.int32))


def test_cluster_vs_oracle(torch_cuda):
    from cpp_projects_amd.search import clusterization, numpy_to_cands

    torch = torch_cuda
    rng = np.random.default_rng(12)
    for trial in range(15):
        n = int(rng.integers(1, 3000 if trial % 3 == 0 else 200))
        c = np.zeros(n, orc.CAND_DTYPE)
        c["mt"] = rng.integers(0, 500, n)
        c["mdt"] = rng.integers(0, 60, n)
        c["mwidth"] = rng.integers(0, 8, n)
        c["msnr"] = rng.integers(0, 20, n).astype(np.float32) * 0.5
        bins = np.array([int(rng.integers(0, 4)), int(rng.integers(0, 3)), int(rng.integers(0, 3))], np.int32)
        order, off, dele, gsz = orc.cluster(c, bins)
        d = numpy_to_cands(c)
        db = torch.from_numpy(bins).cuda()
        gd, gs = clusterization.cluster(d, n, db)
        np.testing.assert_array_equal(gd.cpu().numpy(), dele)
        np.testing.assert_array_equal(gs.cpu().numpy(), gsz)


def test_cluster_golden_image_delegates(torch_cuda):
    from cpp_projects_amd.search import clusterization

    torch = torch_cuda
    img = torch.from_numpy(golden_image()).cuda()
    ok, c, n = clusterization.gather_candidates_in_heap_with_plan(img, 3, 16, thr_tensor(torch, 4.0), 2)
    bins = torch.tensor([2, 1, 2], dtype=torch.int32, device="cuda")
    dele, gsz = clusterization.cluster(c, n, bins)
    assert dele.cpu().tolist() == [1, 7, 4]
    assert gsz.cpu().tolist() == [3, 3, 3]


def test_clusterization_main_writes_log(torch_cuda, tmp_path):
    from cpp_projects_amd.search import clusterization

    torch = torch_cuda
    img = torch.from_numpy(golden_image()).cuda()
    bins = torch.tensor([2, 1, 2], dtype=torch.int32, device="cuda")
    fn = str(tmp_path / "cands.log")
    assert clusterization.clusterization_main(img, 3, 16, thr_tensor(torch, 4.0), 2, bins, fn, 1e-5, 0.0)
    lines = open(fn).read().splitlines()
    assert len(lines) == 4
    assert lines[1].split() == ["1", "4", "4e-05", "0", "1", "140", "3"]
    zero = torch.zeros((3, 16), dtype=torch.float32, device="cuda")
    assert not clusterization.clusterization_main(zero, 3, 16, thr_tensor(torch, 4.0), 2, bins, fn, 1e-5, 0.0)


def test_calc_fdmt_inputs_vs_oracle(torch_cuda):
    from cpp_projects_amd.search import statistical_preprocessing

    torch = torch_cuda
    rng = np.random.default_rng(8)
    R, N = 100, 3000
    x = (rng.standard_normal((R, N)) * 3 + 7).astype(np.float32)
    x[17] *= 9.0
    for mode in (1, 2):
        d = torch.from_numpy(x.copy()).cuda()
        buff = torch.zeros(2 * R + 4, dtype=torch.float32, device="cuda")
        mask = torch.zeros((R, N), dtype=torch.float32, device="cuda")
        statistical_preprocessing.calc_fdmt_inputs(d, R, N, buff, mask, mode)
        torch.cuda.synchronize()
        out, m, (rm, rv, keep, st) = orc.normalise(x, mode)
        np.testing.assert_array_equal(d.cpu().numpy().view(np.uint32), out.view(np.uint32))
        np.testing.assert_array_equal(mask.cpu().numpy(), m)
        b = buff.cpu().numpy()
        np.testing.assert_array_equal(b[:R], rm.astype(np.float32))
        assert b[2 * R] == np.float32(st[0]) and b[2 * R + 1] == np.float32(st[1])
        assert keep[17] == 0


def test_fdmt_normalization_vs_oracle(torch_cuda):
    from cpp_projects_amd.search import fdmt_normalization

    torch = torch_cuda
    rng = np.random.default_rng(3)
    a = rng.standard_normal(10000).astype(np.float32)
    n = rng.integers(0, 50, 10000).astype(np.float32)
    da, dn = torch.from_numpy(a).cuda(), torch.from_numpy(n).cuda()
    out = torch.empty_like(da)
    fdmt_normalization(da, dn, a.size, out)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out.cpu().numpy().view(np.uint32), orc.fdmt_normalization(a, n).view(np.uint32))


def test_detect_max_vs_oracle(torch_cuda):
    from cpp_projects_amd.search import detect_max

    torch = torch_cuda
    rng = np.random.default_rng(9)
    a = rng.standard_normal((6, 900)).astype(np.float32)
    a[4, 300:305] += 3
    n = rng.integers(1, 9, (6, 900)).astype(np.float32)
    got = detect_max(torch.from_numpy(a).cuda(), torch.from_numpy(n).cuda(), 6, 900, 16)
    exp = orc.detect_max(a, n, 16)
    assert got[:3] == exp[:3] and np.float32(got[3]) == np.float32(exp[3])


def test_fused_search_vs_oracle_chain(torch_cuda):
    """normalise -> FDMT -> S/N by FDMT(mask) -> boxcar, against the oracle
    chain run stage by stage (u8 input: bit-exact image and candidates)."""
    import torch

    from cpp_projects_amd.search import SingleBeamSearch, cands_to_numpy

    rng = np.random.default_rng(5)
    nch, ns, dtm = 128, 4096, 127
    x = np.clip(np.round(rng.normal(128, 16, (nch, ns))), 0, 255).astype(np.uint8)
    x[40] = np.clip(np.round(rng.normal(128, 60, ns)), 0, 255).astype(np.uint8)  # RFI row
    s = SingleBeamSearch(1200.0, 1600.0, nch, ns, dtm, in_dtype="u8", box_q=5, threshold=5.0, cand_capacity=1 << 16)
    d_in = torch.from_numpy(x).cuda()
    d_dmt = torch.empty((s.ndt, ns), dtype=torch.float32, device="cuda")
    d_c = torch.empty((1 << 16, 4), dtype=torch.int32, device="cuda")
    d_n = torch.zeros(1, dtype=torch.int64, device="cuda")
    s.run(d_in, d_dmt, d_c, d_n)
    n = s.sort_candidates(d_c, d_n)
    # oracle chain
    xn, mask, (rm, rv, keep, st) = orc.normalise(x, 1)
    assert keep[40] == 0
    p = orc.Plan(1200.0, 1600.0, nch, ns, 1.0, dtm)
    img = p.execute(xn)
    norm = p.execute(mask)
    snr = orc.fdmt_normalization(img, norm)
    np.testing.assert_array_equal(d_dmt.cpu().numpy().view(np.uint32), snr.view(np.uint32))
    ref = orc.boxcar_pow2(snr, 5.0, 5)
    assert n == len(ref)
    np.testing.assert_array_equal(cands_to_numpy(d_c, n).view(np.int32), ref.view(np.int32))
