"""Oracle correlate / findlocalmaxima / matched_filter (src/correlation.jl,
src/detection.jl), pinned on the CPU:

* correlate against scipy.ndimage.correlate (mode 'nearest' is the
  reference's :replicate, 'constant' with cval 0 its :zero; both centre an
  odd kernel on its middle sample) and the reference's spot checks
  (test/correlation_tests.jl:146-215);
* findlocalmaxima against a brute-force strict 3 x 3 maximum and the
  reference's cases (test/detection_tests.jl:450-529);
* matched_filter against closed forms (zero-sum kernel, unbiased flux on a
  noiseless source, zero significance on a constant image, the N(0, 1)
  calibration) and the reference's cases (test/detection_tests.jl:1-697).
"""
import math

import numpy as np
import pytest
from scipy import ndimage

import oracle_detect as od

rng0 = np.random.default_rng(7)


def _gk(k, s):
    x = np.linspace(-(k // 2), k // 2, k)
    g = np.exp(-0.5 * (x / s) ** 2)
    g /= g.sum()
    return np.outer(g, g)


def _place(img, y, x, flux, s):
    """Gaussian source at 1-based (y, x) (the reference's _place_source!)."""
    Y, X = img.shape
    k = math.ceil(8 * s)
    k = k if k % 2 else k + 1
    for j in range(max(1, round(x - k // 2)), min(X, round(x + k // 2)) + 1):
        for i in range(max(1, round(y - k // 2)), min(Y, round(y + k // 2)) + 1):
            img[i - 1, j - 1] += flux * math.exp(-((j - x) ** 2 + (i - y) ** 2) / (2 * s * s)) / (2 * math.pi * s * s)
    return img


@pytest.mark.parametrize("shape", [(40, 37), (9, 9), (3, 50), (5, 4)])
@pytest.mark.parametrize("ks", [(1, 1), (3, 3), (5, 7), (11, 11), (13, 3)])
@pytest.mark.parametrize("border", ["replicate", "zero"])
def test_correlate_matches_scipy(shape, ks, border):
    img = rng0.standard_normal(shape)
    K = rng0.standard_normal(ks)
    got = od.correlate(img, K, border)
    want = ndimage.correlate(img, K, mode="nearest" if border == "replicate" else "constant", cval=0.0)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)


def test_correlate_spot_checks_and_oversized_kernel():
    kern = np.array([[1, 2, 1], [2, 4, 2], [1, 2, 1]], float) / 16
    img = np.zeros((9, 9))
    img[4, 4] = 1.0
    out = od.correlate(img, kern, "zero")
    np.testing.assert_allclose(out[3:6, 3:6], kern)
    assert np.all(out[:3] == 0) and np.all(out[6:] == 0)
    diag = np.eye(3) / 3
    np.testing.assert_allclose(od.correlate(img, diag, "zero")[3:6, 3:6], diag)
    # a shift kernel moves the image: correlation reads img[row + i - r - 1, ...]
    shift = np.zeros((3, 3))
    shift[2, 1] = 1.0
    a = rng0.standard_normal((7, 6))
    np.testing.assert_array_equal(od.correlate(a, shift, "zero")[:-1], a[1:])
    # kernel larger than the image
    stamp = rng0.standard_normal((4, 5))
    K = rng0.standard_normal((9, 11))
    np.testing.assert_allclose(od.correlate(stamp, K, "replicate"), ndimage.correlate(stamp, K, mode="nearest"),
                               rtol=1e-12)
    with pytest.raises(ValueError, match="odd"):
        od.correlate(stamp, np.ones((4, 4)))


def _brute_maxima(a, edges):
    H, W = a.shape
    out = []
    for c in range(W):
        for r in range(H):
            if not edges and not (0 < r < H - 1 and 0 < c < W - 1):
                continue
            nb = [a[rr, cc] for rr in range(r - 1, r + 2) for cc in range(c - 1, c + 2)
                  if (rr, cc) != (r, c) and 0 <= rr < H and 0 <= cc < W]
            if all(a[r, c] > v for v in nb):
                out.append((r + 1, c + 1))
    return out


def _order_key(p, H, W):
    r, c = p
    if 1 < r < H and 1 < c < W:
        return (0, c, r)
    if 1 < c < W:
        return (1, c, 0 if r == 1 else 1)
    if 1 < r < H:
        return (2, r, 0 if c == 1 else 1)
    return (3, (0 if r == 1 else 2) + (0 if c == 1 else 1), 0)


@pytest.mark.parametrize("shape", [(30, 25), (3, 3), (2, 7), (7, 2), (4, 4)])
def test_findlocalmaxima_matches_brute_force(shape):
    for seed in range(4):
        a = np.random.default_rng(seed).integers(0, 6, shape).astype(float)   # ties are common
        H, W = shape
        for edges in (False, True):
            got = od.findlocalmaxima(a, edges)
            want = sorted(_brute_maxima(a, edges), key=lambda p: _order_key(p, H, W))
            assert got == want, (seed, edges)


def test_findlocalmaxima_reference_cases():
    img = np.zeros((11, 11))
    img[5, 6] = 1.0
    assert od.findlocalmaxima(img) == [(6, 7)]
    img = np.zeros((20, 20))
    img[4, 4] = img[14, 9] = img[9, 16] = 1.0
    assert sorted(od.findlocalmaxima(img)) == [(5, 5), (10, 17), (15, 10)]
    assert od.findlocalmaxima(np.ones((10, 10))) == []
    img = np.zeros((9, 9))
    img[4, 4], img[4, 5] = 2.0, 1.0
    assert od.findlocalmaxima(img) == [(5, 5)]
    img = np.zeros((8, 8))
    img[0, 3] = 1.0
    assert od.findlocalmaxima(img, edges=True) == [(1, 4)] and od.findlocalmaxima(img, edges=False) == []
    img = np.zeros((6, 6))
    img[0, 0] = img[5, 5] = 1.0
    assert od.findlocalmaxima(img) == [(1, 1), (6, 6)]
    assert od.findlocalmaxima(np.zeros((1, 1))) == [] and od.findlocalmaxima(np.array([[1.0, 0], [0, 0]])) == []
    img = np.zeros((9, 9))
    img[4, 4] = img[4, 5] = 1.0
    assert od.findlocalmaxima(img) == []
    with pytest.raises(ValueError):
        od.findlocalmaxima(np.arange(5.0).reshape(1, 5), edges=True)


def test_matched_filter_kernel_normalisation_and_unbiased_flux():
    kern = _gk(11, 2.0)
    r = od.matched_filter(np.zeros((100, 100)), kern, normalize_zerosum=True)
    assert abs(r.kernel.sum()) < 1e-13
    N = kern.size
    denom = (kern ** 2).sum() - kern.sum() ** 2 / N
    assert r.kernel_norm == pytest.approx(math.sqrt(denom), rel=1e-14)
    np.testing.assert_allclose(r.kernel, (kern - kern.mean()) / denom, rtol=1e-12, atol=1e-18)
    r = od.matched_filter(np.zeros((100, 100)), kern, normalize_zerosum=False)
    assert r.kernel.sum() > 0 and r.kernel_norm == pytest.approx(math.sqrt((kern ** 2).sum()), rel=1e-14)
    # noiseless source on a flat background: the zero-sum flux estimate is F
    # when the source sits on a pixel centre (E[corr(F P, K)] = F exactly)
    k9 = _gk(9, 1.5)
    img = np.full((40, 40), 10.0)
    img[15:24, 15:24] += 50.0 * k9
    r = od.matched_filter(img, k9, normalize_zerosum=True)
    i = r.peaks.index((20, 20))
    assert r.peak_fluxes[i] == pytest.approx(50.0, rel=1e-12)
    # and the reference's analytic source (test/detection_tests.jl:117-129)
    img = _place(np.full((40, 40), 10.0), 20.0, 20.0, 50.0, 1.5)
    r = od.matched_filter(img, k9, normalize_zerosum=True)
    d = [(p[0] - 20) ** 2 + (p[1] - 20) ** 2 for p in r.peaks]
    assert r.peak_fluxes[int(np.argmin(d))] == pytest.approx(50.0, rel=0.05)


def test_matched_filter_background_and_calibration():
    k9 = _gk(9, 1.5)
    img = np.full((40, 40), 100.0)
    zs = od.matched_filter(img, k9, normalize_zerosum=True)
    nzs = od.matched_filter(img, k9, normalize_zerosum=False)
    assert np.max(np.abs(zs.significance_map)) < 1e-12 and np.max(np.abs(nzs.significance_map)) > 100
    # N(0, 1) significance on pure unit noise with inv_var = 1
    rng = np.random.default_rng(303)
    sig = []
    for _ in range(20):
        r = od.matched_filter(rng.standard_normal((50, 50)), _gk(11, 2.0), inv_var=np.ones((50, 50)), sigma=0.0)
        sig.append(r.significance_map[9:40, 9:40].ravel())
    sig = np.concatenate(sig)
    assert abs(sig.std() - 1) < 0.15 and abs(sig.mean()) < 0.05
    # with inv_var the significance is num / sqrt(den) exactly
    img = rng.standard_normal((30, 30))
    iv = rng.uniform(0.5, 2.0, (30, 30))
    r = od.matched_filter(img, _gk(7, 1.2), inv_var=iv, sigma=0.0)
    K = r.kernel
    num = od.correlate(iv * img, K)
    den = od.correlate(iv, K ** 2)
    np.testing.assert_allclose(r.significance_map, num / np.sqrt(den), rtol=1e-12)
    np.testing.assert_allclose(r.smoothed_inv_var, den, rtol=1e-12)


def test_matched_filter_inv_var_masking_and_edge_cases():
    rng = np.random.default_rng(9)
    k = _gk(9, 1.5)
    img = _place(rng.standard_normal((40, 40)) * 0.3, 20.0, 20.0, 40.0, 1.5)
    iv = np.full(img.shape, 1 / 0.09)
    a = od.matched_filter(img, k, inv_var=iv)
    iv2 = iv.copy()
    iv2[5, 30] = np.nan
    img2 = img.copy()
    img2[5, 30] = np.nan
    b = od.matched_filter(img2, k, inv_var=iv2)
    assert np.all(np.isfinite(b.significance_map))
    assert (20, 20) in a.peaks and (20, 20) in b.peaks
    # zero weights give zero significance
    r = od.matched_filter(img, k, inv_var=np.zeros_like(img), sigma=0.0)
    assert np.all(r.significance_map == 0) and r.peaks == []
    # degenerate kernels: uniform (zero-sum denominator 0) and zero
    r = od.matched_filter(img, np.ones((5, 5)))
    assert r.kernel_norm == 1.0 and np.all(r.kernel == 1.0)
    r = od.matched_filter(img, np.zeros((5, 5)), normalize_zerosum=False)
    assert r.kernel_norm == 1.0
    # 1 x 1 kernel, sigma = 0 keeps every interior local maximum
    r = od.matched_filter(img, np.ones((1, 1)), normalize_zerosum=False, sigma=-np.inf)
    assert r.peaks == od.findlocalmaxima(r.significance_map, edges=False)
    with pytest.raises(ValueError, match="odd"):
        od.matched_filter(img, np.ones((4, 4)))
