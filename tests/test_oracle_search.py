"""Pins the oracle's normalisation, boxcar gathers, clustering and detector.

Golden: the reference's Candidates1.csv / Candidates2.csv (input from
clusterization_forpravir/clusterization_test.cu:63-81); KATs from
regrouping_plan/kernel.cu:100-132,211, select_delegates/select_delegates.cu:
332-352 and final_sort/kernel.cu:37-55; brute force and closed forms for the
rest.
"""
import csv
import os

import numpy as np
import pytest

from oracle import orc

GD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden_image():
    img = np.full((3, 16), 2.0, np.float32)
    img[0, 4] = 140.0
    img[0, 11] = 120.0
    img[2, 4] = 100.0
    return img


def read_csv(fn):
    with open(os.path.join(GD, fn)) as f:
        return [(int(r["mt"]), int(r["mdt"]), int(r["mwidth"]), float(r["msnr"])) for r in csv.DictReader(f)]


def as_tuples(c):
    return [(int(a["mt"]), int(a["mdt"]), int(a["mwidth"]), float(a["msnr"])) for a in c]


def test_boxcar_all_widths_matches_candidates2_in_order():
    got = as_tuples(orc.boxcar_all(golden_image(), 4.0, 2))
    exp = read_csv("Candidates2.csv")
    assert [g[:3] for g in got] == [e[:3] for e in exp]
    for g, e in zip(got, exp):
        assert g[3] == pytest.approx(e[3], rel=1e-5)  # csv holds 6 significant digits


def test_boxcar_pow2_matches_candidates1_as_set():
    got = sorted(as_tuples(orc.boxcar_pow2(golden_image(), 4.0, 1)))
    exp = sorted(read_csv("Candidates1.csv"))
    assert [g[:3] for g in got] == [e[:3] for e in exp]
    for g, e in zip(got, exp):
        assert g[3] == pytest.approx(e[3], rel=1e-5)


def test_boxcar_closed_form_values():
    # (2 + 140) / sqrt(2) etc. computed in fp32 as the reference does
    c = orc.boxcar_all(golden_image(), 4.0, 2)
    s = {(int(a["mt"]), int(a["mdt"]), int(a["mwidth"])): a["msnr"] for a in c}
    assert s[(3, 0, 2)] == np.float32(142.0) / np.sqrt(np.float32(2.0))
    assert s[(4, 0, 1)] == np.float32(140.0)


def brute_boxcar(img, thr, widths, log2=False):
    out = []
    R, N = img.shape
    for r in range(R):
        for t in range(N):
            for w in widths:
                if t + w > N:
                    continue
                s = np.float32(0.0)
                for j in range(w):
                    s = np.float32(s + img[r, t + j])
                v = np.float32(s / np.sqrt(np.float32(w)))
                if v > np.float32(thr):
                    out.append((t, r, int(np.log2(w)) if log2 else w, float(v)))
    return out


def test_boxcar_brute_force_tiny():
    rng = np.random.default_rng(5)
    img = rng.standard_normal((4, 40)).astype(np.float32)
    img[1, 7:11] += 3
    for q in (0, 1, 2, 3):
        got = as_tuples(orc.boxcar_pow2(img, 1.5, q))
        exp = brute_boxcar(img, 1.5, [1 << k for k in range(q + 1)], log2=True)
        assert got == exp
    got = as_tuples(orc.boxcar_all(img, 1.5, 5))
    assert got == brute_boxcar(img, 1.5, [1, 2, 3, 4, 5])


def cands(rows):
    c = np.zeros(len(rows), orc.CAND_DTYPE)
    for i, r in enumerate(rows):
        c[i] = r
    return c


def test_regrouping_kat():
    """regrouping_plan/kernel.cu: two groups over mt with gap 2 -> {4, 4} subgroups."""
    rows = [(10, 2, 3, 4.0), (26, 5, 6, 7.0), (12, 8, 9, 10.0), (50, 11, 12, 13.0), (28, 14, 15, 16.0),
            (11, 3, 4, 5.0), (27, 6, 7, 8.0), (25, 9, 10, 11.0), (51, 12, 13, 14.0), (29, 15, 16, 17.0),
            (52, 18, 19, 20.0), (70, 21, 22, 23.0), (121, 39, 40, 41.0), (122, 42, 43, 44.0),
            (80, 24, 25, 26.0), (90, 27, 28, 29.0), (91, 30, 31, 32.0), (92, 33, 34, 35.0),
            (120, 36, 37, 38.0), (150, 45, 46, 47.0)]
    c = cands(rows)
    g0 = c[:12]
    g1 = c[12:]
    # mt-only split counts (huge bins on the other axes)
    for g, exp in ((g0, 4), (g1, 4)):
        _, off, _, _ = orc.cluster(g, [2, 10**9, 10**9])
        assert len(off) - 1 == exp


KAT_SNR = [1.0, 2.0, 3.0, 0.5, 2.5, 1.5, 4.0, 3.5, 0.7, 1.3, 5.0, 3.3, 2.7, 4.5, 3.2]
KAT_PERM = [4, 6, 12, 7, 0, 2, 1, 8, 5, 14, 11, 9, 10, 3, 13]


def test_select_delegates_kat():
    """select_delegates/select_delegates.cu:323-347: 15 candidates, groups
    beginning at {0, 3, 8} of a permuted index array -> delegates {6, 7, 10}."""
    c = cands([(i + 1, i + 2, i + 3, v) for i, v in enumerate(KAT_SNR)])
    assert orc.select_delegates(c, KAT_PERM, [0, 3, 8]).tolist() == [6, 7, 10]
    # the same groups made by the three-key clustering: members of group g
    # share mt = 100 (g + 1), groups are 100 apart, the other keys are equal
    c2 = cands([(0, 0, 0, v) for v in KAT_SNR])
    for g, (b, e) in enumerate(((0, 3), (3, 8), (8, 15))):
        for j in KAT_PERM[b:e]:
            c2[j]["mt"] = 100 * (g + 1)
    _, off, dele, gsz = orc.cluster(c2, [2, 1, 2])
    assert dele.tolist() == [6, 7, 10] and gsz.tolist() == [3, 5, 7]


def test_select_delegates_ties_and_nan():
    """thrust::max_element's sequential rule: the first maximum in group
    order wins; a NaN only where it is a group's first element."""
    c = cands([(0, 0, 0, v) for v in (3.0, 5.0, 5.0, float("nan"), 1.0, float("nan"), 2.0)])
    d = orc.select_delegates(c, [2, 1, 0, 3, 4, 6, 5], [0, 3, 5])
    assert d.tolist() == [2, 3, 6]
    assert orc.select_delegates(c, [0, 1], [0, 2]).tolist() == [1, -1]


def test_final_sort_kat():
    """final_sort/kernel.cu:37-55: delegates {8, 6, 5} of candidates with
    mt = [10, 30, 20, 50, 40, 15, 35, 25, 55, 45] sorted by mt -> {5, 6, 8};
    here through the clustering, each candidate its own group."""
    mts = [10, 30, 20, 50, 40, 15, 35, 25, 55, 45]
    picked = [8, 6, 5]
    c = cands([(mts[j], 3 * k, 0, 1.0) for k, j in enumerate(picked)])
    _, off, dele, gsz = orc.cluster(c, [0, 0, 0])
    assert [picked[int(i)] for i in dele] == [5, 6, 8]
    assert gsz.tolist() == [1, 1, 1]


def brute_cluster(c, bins):
    groups = [list(range(len(c)))]
    for k, key in enumerate(("mt", "mdt", "mwidth")):
        ng = []
        for g in groups:
            g = sorted(g, key=lambda j: int(c[j][key]))
            cur = [g[0]]
            for a, b in zip(g, g[1:]):
                if int(c[b][key]) - int(c[a][key]) > bins[k]:
                    ng.append(cur)
                    cur = []
                cur.append(b)
            ng.append(cur)
        groups = ng
    return sorted(sorted(g) for g in groups)


def test_cluster_matches_brute_force():
    rng = np.random.default_rng(11)
    for trial in range(20):
        n = int(rng.integers(1, 60))
        rows = [(int(rng.integers(0, 40)), int(rng.integers(0, 10)), int(rng.integers(0, 4)),
                 float(rng.integers(0, 6))) for _ in range(n)]
        c = cands(rows)
        bins = [int(rng.integers(0, 4)), int(rng.integers(0, 3)), int(rng.integers(0, 3))]
        order, off, dele, gsz = orc.cluster(c, bins)
        groups = sorted(sorted(order[off[g]:off[g + 1]].tolist()) for g in range(len(off) - 1))
        assert groups == brute_cluster(c, bins)
        # delegates: max msnr, ties -> lowest index; sorted by mt (stable)
        members = {int(order[i]): g for g in range(len(off) - 1) for i in range(off[g], off[g + 1])}
        for d, sz in zip(dele, gsz):
            grp = [j for j in range(n) if members[j] == members[int(d)]]
            best = max(grp, key=lambda j: (c[j]["msnr"], -j))
            assert int(d) == best
            assert sz == len(grp)
        assert all(c[dele[i]]["mt"] <= c[dele[i + 1]]["mt"] for i in range(len(dele) - 1))


def test_cluster_golden_image():
    c = orc.boxcar_all(golden_image(), 4.0, 2)
    order, off, dele, gsz = orc.cluster(c, [2, 1, 2])
    assert dele.tolist() == [1, 7, 4]
    assert gsz.tolist() == [3, 3, 3]


def test_stats_closed_form_and_cleaning():
    R, N = 64, 4096
    rng = np.random.default_rng(2)
    x = rng.standard_normal((R, N)).astype(np.float32) * 2 + 5
    x[10] = rng.standard_normal(N).astype(np.float32) * 10 + 5  # variance 100 -> dropped
    rm, rv, keep, st = orc.stats(x)
    np.testing.assert_allclose(rm, x.astype(np.float64).mean(1), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(rv, x.astype(np.float64).var(1), rtol=1e-9)
    assert keep[10] == 0 and keep.sum() == R - 1
    allx = x.astype(np.float64)
    assert st[0] == pytest.approx(allx.mean(), rel=1e-12)
    assert st[1] == pytest.approx(allx.std(), rel=1e-9)
    out, mask, _ = orc.normalise(x)
    assert (out[10] == 0).all() and (mask[10] == 0).all() and (mask[0] == 1).all()
    mu, sg = np.float32(st[0]), np.float32(st[1])
    np.testing.assert_array_equal(out[3], (x[3] - mu) / sg)


def test_stats_integer_input_exact():
    R, N = 16, 1000
    x = np.random.default_rng(4).integers(0, 256, (R, N)).astype(np.uint8)
    rm, rv, keep, st = orc.stats(x)
    s1 = x.astype(np.int64).sum(1)
    s2 = (x.astype(np.int64) ** 2).sum(1)
    np.testing.assert_array_equal(rm, s1 / N)
    np.testing.assert_array_equal(rv, s2 / N - (s1 / N) ** 2)
    # uint8 and its float32 promotion give identical statistics
    rm2, rv2, _, st2 = orc.stats(x.astype(np.float32))
    np.testing.assert_array_equal(rm, rm2)
    np.testing.assert_array_equal(st, st2)


def variance_rows(var, n=512):
    """Rows of n samples alternating +a, -a: mean 0, variance exactly a*a in FP64."""
    a = np.sqrt(np.asarray(var, np.float64)).astype(np.float32)
    x = np.empty((len(a), n), np.float32)
    x[:, 0::2] = a[:, None]
    x[:, 1::2] = -a[:, None]
    return x, a.astype(np.float64) ** 2


def cleaning_ratio(v):
    v = np.asarray(v, np.float64)
    return np.abs(v[-1] - v.mean()) / v.std()


@pytest.mark.parametrize("target,kept", [(3.99, True), (4.01, False), (3.5, True)])
def test_stats_cleaning_at_the_4_sigma_boundary(target, kept):
    """A row whose variance lies just inside 4 sigma of the mean row variance
    is kept (a 3-sigma rule would drop it) and just outside is dropped
    (Statistical_preprocessing.cu:521-537, strict >)."""
    base = [1.0 + 0.05 * (i % 7) for i in range(63)]
    lo, hi = 1.0, 1e6
    for _ in range(200):  # the outlier variance giving the target ratio
        mid = 0.5 * (lo + hi)
        if cleaning_ratio(base + [mid]) < target:
            lo = mid
        else:
            hi = mid
    x, v = variance_rows(base + [0.5 * (lo + hi)])
    r = cleaning_ratio(v)
    assert abs(r - target) < 1e-3 and (r > 3.0)
    rm, rv, keep, st = orc.stats(x)
    np.testing.assert_array_equal(rv, v)
    assert bool(keep[-1]) == kept and keep[:-1].all()


def test_per_channel_mode():
    x = np.random.default_rng(6).standard_normal((8, 512)).astype(np.float32) * np.arange(1, 9, dtype=np.float32)[:, None]
    out, mask, (rm, rv, keep, st) = orc.normalise(x, mode=2)
    np.testing.assert_allclose(out.std(1), 1.0, rtol=1e-5)


def test_fdmt_normalization_closed_form():
    a = np.array([1.0, 4.0, -3.0, 0.0], np.float32)
    n = np.array([4.0, 16.0, 0.0, 1.0], np.float32)
    o = orc.fdmt_normalization(a, n)
    assert o[0] == 0.5 and o[1] == 1.0 and o[3] == 0.0
    assert o[2] == np.float32(-3.0) / np.sqrt(np.float32(1e-8))


def test_detect_max_brute():
    rng = np.random.default_rng(9)
    a = rng.standard_normal((5, 50)).astype(np.float32)
    a[2, 20:24] += 4
    n = np.full_like(a, 4.0)
    w, r, t, s = orc.detect_max(a, n, 8)
    best = None
    for ww in range(1, 9):
        for rr in range(5):
            for tt in range(50 - ww + 1):
                v = np.float32(0)
                for j in range(ww):
                    v = np.float32(v + np.float32(a[rr, tt + j] / np.sqrt(np.float32(4.0 + 1e-7))))
                v = np.float32(v / np.sqrt(np.float32(ww)))
                if best is None or v > best[0]:
                    best = (v, ww, rr, tt)
    assert (w, r, t) == best[1:] and s == best[0]


def detector_kat_input():
    rows, cols = 256, (1 << 20) // 256
    a = np.ones((rows, cols), np.float32)
    n = np.ones((rows, cols), np.float32)
    f = a.reshape(-1)
    f[3], f[12] = 20, 40
    f[[8, 9, 11]] = 0
    f[13:18] = 0
    n.reshape(-1)[cols + 3:cols + 5] = 0.001
    return a, n


def test_tiled_detectors_on_the_reference_input():
    """detection/kernel.cu:645-676 with WndWidth 10, blocks of 256: the
    boxcar S/N maximum is the width-2 window over the two near-zero-norm
    cells of row 1, (1/sqrt(1e-7 + 0.001)) * 2 / sqrt(2); the energy ratio is
    maximal on those cells too: 1 / (n*n) with the n*n sums of that tile
    differing by 4 ulp of 3.0 (2^-20), so sqrt(2^20) = 1024 exactly."""
    a, n = detector_kat_input()
    w, r, c, v = orc.detect_tiles(1, a, n, 10, 256)
    assert (w, r, c) == (2, 1, 3)
    y = 1.0 / np.sqrt(1e-7 + np.float64(np.float32(0.001)))
    assert v == pytest.approx(2 * y / np.sqrt(2.0), rel=1e-6)
    w, r, c, v = orc.detect_tiles(2, a, n, 10, 256)
    assert (w, r, c, v) == (1, 1, 3, 1024.0)
    # the whole-row detector finds the same window as the tiled one here
    assert orc.detect_max(a, n, 10)[:3] == (2, 1, 3)


def test_tiled_detectors_vs_fp64_windows():
    """Brute force: every tile's best window value is within the
    prefix-difference rounding bound of its FP64 window sum, and windows that
    cross a tile never win in mode 1."""
    rng = np.random.default_rng(31)
    for trial in range(6):
        R, N, W, tile = 3, int(rng.integers(40, 300)), int(rng.integers(1, 12)), int(rng.choice([16, 32, 64]))
        a = rng.standard_normal((R, N)).astype(np.float32)
        a[1, N // 2:N // 2 + 5] += 3.0
        n = rng.uniform(0.5, 2.0, (R, N)).astype(np.float32)
        for mode in (1, 2):
            w, r, c, v = orc.detect_tiles(mode, a, n, W, tile)
            t0 = (c // tile) * tile
            assert c + w <= N and (mode == 2 or c + w <= t0 + tile)
            if mode == 1:
                y = a[r].astype(np.float64) / np.sqrt(1e-7 + n[r].astype(np.float64))
                exact = y[c:c + w].sum() / np.sqrt(w)
                scale = np.abs(y[t0:c + w]).sum() / np.sqrt(w)
            else:
                x2 = a[r].astype(np.float64) ** 2
                n2 = n[r].astype(np.float64) ** 2
                exact = np.sqrt(x2[c:c + w].sum() / n2[c:c + w].sum())
                scale = exact * (1 + (x2[t0:c + w].sum() / x2[c:c + w].sum() + n2[t0:c + w].sum() /
                                      n2[c:c + w].sum()))
            assert abs(v - exact) <= 2e-6 * (N + W) * scale + 1e-6
            # no window of the same row scores more than the winner by more than the bound
            if mode == 1:
                for t in range(N):
                    tt = (t // tile) * tile
                    for ww in range(1, W + 1):
                        if t + ww <= min(N, tt + tile):
                            e2 = y[t:t + ww].sum() / np.sqrt(ww)
                            assert e2 <= exact + 4e-6 * (N + W) * (np.abs(y).sum() / np.sqrt(ww)) + 1e-5
