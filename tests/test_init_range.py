"""The fused init's per-channel range test for 8-bit input
(u8_init_plain in cpp_projects_amd/csrc/fdmt_kernels.hip): for v = 0..255,
a(v) = fl(v - mu) and x(v) = fl(a / sigma) in float32 are non-decreasing in v
and vanish only at v = mu, so five samples (0, 255 and the integers next to mu)
decide whether every nonzero |a| and |x| lies in [2^-60, 2^60].  Checked here,
in IEEE float32 arithmetic on the CPU, against all 256 samples for means and
deviations across and at the edges of the range: the five-sample test never
passes a channel that the exhaustive test rejects, and for |mu| <= 2^30 the two
agree."""
import numpy as np

LO, HI = np.float32(2.0 ** -60), np.float32(2.0 ** 60)


def five_sample_test(mu, sg):
    """The kernel's decision, restated."""
    mu, sg = np.float32(mu), np.float32(sg)
    if not (LO <= sg <= HI) or not (abs(mu) <= np.float32(2.0 ** 30)):
        return False
    f = int(np.floor(mu))
    ok = True
    for v in (0, 255, f - 1, f, f + 1):
        if v < 0 or v > 255:
            continue
        a = np.float32(np.float32(v) - mu)
        if a == 0:
            continue
        x = np.float32(a / sg)
        ok = ok and LO <= abs(a) <= HI and LO <= abs(x) <= HI
    return bool(ok)


def exhaustive_test(mu, sg):
    mu, sg = np.float32(mu), np.float32(sg)
    if not (LO <= sg <= HI):
        return False
    v = np.arange(256, dtype=np.float32)
    with np.errstate(all="ignore"):
        a = (v - mu).astype(np.float32)
        x = (a / sg).astype(np.float32)
    nz = a != 0
    aa, ax = np.abs(a[nz]), np.abs(x[nz])
    return bool(np.all((aa >= LO) & (aa <= HI) & (ax >= LO) & (ax <= HI)))


def cases():
    rng = np.random.default_rng(12)
    mus = list(rng.uniform(-10, 270, 300)) + [0.0, 255.0, 127.5, 1e-30, -1e-30, 2.0 ** -70, 2.0 ** -59, 256.0,
                                              -1.0, 2.0 ** 30, 2.0 ** 31, float("nan"), float("inf")]
    for k in (0, 1, 17, 128, 254, 255):  # integers and their float32 neighbours
        for d in (-3, -1, 1, 3):
            mus.append(float(np.nextafter(np.float32(k), np.float32(k + d), dtype=np.float32)))
    sgs = list(np.exp(rng.uniform(np.log(1e-3), np.log(1e3), 40))) + [
        2.0 ** -61, 2.0 ** -60, 2.0 ** -53, 2.0 ** -40, 2.0 ** 40, 2.0 ** 53, 2.0 ** 59, 2.0 ** 60, 2.0 ** 61,
        0.0, -1.0, float("nan"), float("inf")]
    for mu in mus:
        for sg in sgs:
            yield mu, sg


def test_five_samples_decide_all_256():
    passed = 0
    for mu, sg in cases():
        f, e = five_sample_test(mu, sg), exhaustive_test(mu, sg)
        assert not f or e, (mu, sg)  # never passes a channel the full test rejects
        if abs(np.float32(mu)) <= np.float32(2.0 ** 30):
            assert f == e, (mu, sg)
        passed += f
    assert passed > 1000  # typical channels pass


def test_samples_are_monotone():
    """The premise: a(v) and x(v) never decrease as v grows, and x(v) is
    never -0 (a = +0 exactly at v = mu, and a nonzero a stays nonzero in
    range)."""
    rng = np.random.default_rng(3)
    v = np.arange(256, dtype=np.float32)
    for _ in range(500):
        mu = np.float32(rng.uniform(-5, 260))
        sg = np.float32(np.exp(rng.uniform(-20, 20)))
        a = (v - mu).astype(np.float32)
        x = (a / sg).astype(np.float32)
        assert np.all(np.diff(a) >= 0) and np.all(np.diff(x) >= 0)
        assert not np.any((x == 0) & np.signbit(x))
