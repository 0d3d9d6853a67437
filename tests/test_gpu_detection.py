"""GPU parity of the detection step (correlate, findlocalmaxima,
matched_filter; src/correlation.jl, src/detection.jl) against the FP64
oracle through the C ABI.  Bars: peak lists bit-exact; correlations and
significances within float64 summation-order rounding (1e-11 relative to the
map's scale; the factored path of a rank-1 kernel differs from the direct
sum by rounding only, as in the reference)."""
import numpy as np
import pytest

import oracle_detect as od
from _support import load_pkg

pkg = load_pkg()
sim = pkg.simulation

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return pkg.Context(0)


def _gk(k, s):
    x = np.linspace(-(k // 2), k // 2, k)
    g = np.exp(-0.5 * (x / s) ** 2)
    g /= g.sum()
    return np.outer(g, g)


def _close(a, b, tol=1e-11):
    scale = max(np.max(np.abs(b)), 1e-300)
    np.testing.assert_allclose(a, b, rtol=0, atol=tol * scale)


@pytest.mark.parametrize("border", ["replicate", "zero"])
def test_correlate_paths_match_oracle(ctx, border):
    rng = np.random.default_rng(1)
    for shape in [(64, 48), (17, 5), (3, 40), (100, 130)]:
        img = rng.standard_normal(shape)
        for K in (rng.standard_normal((5, 7)),       # 2-D, window in LDS
                  rng.standard_normal((71, 69)),     # 2-D, window through the cache
                  _gk(15, 2.0),                      # rank 1: two 1-D passes
                  np.ones((1, 1)) * 0.5, rng.standard_normal((1, 9)), rng.standard_normal((9, 1))):
            _close(pkg.correlate(img, K, border, ctx=ctx), od.correlate(img, K, border))
    with pytest.raises(pkg.ArgumentError, match="odd"):
        pkg.correlate(img, np.ones((4, 3)), ctx=ctx)
    with pytest.raises(pkg.ArgumentError, match="border"):
        pkg.correlate(img, np.ones((3, 3)), "circular", ctx=ctx)


def test_findlocalmaxima_bit_exact(ctx):
    for seed, shape in enumerate([(50, 40), (3, 3), (2, 9), (9, 2), (128, 7), (1, 2)]):
        a = np.random.default_rng(seed).integers(0, 5, shape).astype(float)
        for edges in (True, False):
            assert pkg.findlocalmaxima(a, edges=edges, ctx=ctx) == od.findlocalmaxima(a, edges), (shape, edges)
    big = np.random.default_rng(9).standard_normal((700, 650))
    assert pkg.findlocalmaxima(big, ctx=ctx) == od.findlocalmaxima(big)
    with pytest.raises(pkg.ArgumentError):
        pkg.findlocalmaxima(np.arange(5.0).reshape(1, 5), edges=True, ctx=ctx)


@pytest.mark.parametrize("zerosum", [True, False])
@pytest.mark.parametrize("weighted", [True, False])
def test_matched_filter_matches_oracle(ctx, zerosum, weighted):
    rng = np.random.default_rng(3)
    ny, nx = 180, 150
    y, x, f = sim.simulate_sources(5, 60, (3.0, ny - 2.0), (3.0, nx - 2.0), flux=(50.0, 3000.0))
    img = sim.render_sources(pkg.CircularGaussianPSF(fwhm=2.5), (ny, nx), y, x, f, 12)
    img = (img + 10.0 + rng.standard_normal((ny, nx))).astype(np.float32).astype(np.float64)
    iv = None
    if weighted:
        iv = rng.uniform(0.5, 1.5, (ny, nx)).astype(np.float32).astype(np.float64)
        iv[20:23, 40:44] = 0.0
        iv[100, 100] = np.nan
        img[100, 100] = np.nan
    kern = _gk(13, 1.1)
    g = pkg.matched_filter(img, kern, inv_var=iv, normalize_zerosum=zerosum, sigma=3.0, ctx=ctx)
    o = od.matched_filter(img, kern, inv_var=iv, normalize_zerosum=zerosum, sigma=3.0)
    assert len(o.peaks) > 10
    assert g.peaks == o.peaks
    np.testing.assert_allclose(g.kernel, o.kernel, rtol=1e-15, atol=0)
    assert g.kernel_norm == o.kernel_norm
    _close(g.significance_map, o.significance_map)
    np.testing.assert_allclose(g.peak_significances, o.peak_significances, rtol=1e-10)
    np.testing.assert_allclose(g.peak_fluxes, o.peak_fluxes, rtol=1e-10)
    if weighted:
        _close(g.smoothed_inv_var, o.smoothed_inv_var)
    else:
        _close(g.smoothed_image, o.smoothed_image)


def test_matched_filter_convenience_and_catalogue(ctx):
    """fwhm / model kernels (detection.jl:207-232) and the detected catalogue
    feeding fit_all_stars (_extract_source_catalog for a MatchedFilterResult)."""
    ny, nx = 120, 110
    ys = np.array([30.0, 60.0, 90.0, 40.0])
    xs = np.array([25.0, 70.0, 40.0, 90.0])
    fl = np.array([2000.0, 1500.0, 2500.0, 1800.0])
    psf = pkg.CircularGaussianPSF(fwhm=2.5)
    img = sim.render_sources(psf, (ny, nx), ys, xs, fl, 12) + 5.0
    img += np.random.default_rng(4).standard_normal(img.shape) * 0.5
    r1 = pkg.matched_filter(img, 2.5, sigma=10.0, ctx=ctx)
    r2 = pkg.matched_filter(img, (2.5, 2.5), sigma=10.0, ctx=ctx)
    r3 = pkg.matched_filter(img, psf, sigma=10.0, ctx=ctx)
    assert r1.kernel.shape == (15, 15) and r3.kernel.shape == (15, 15)
    np.testing.assert_allclose(r1.kernel, r2.kernel, rtol=1e-12)
    for r in (r1, r3):
        assert sorted(r.peaks) == sorted(zip(ys.astype(int).tolist(), xs.astype(int).tolist()))
    kern = pkg.render(psf)
    o = od.matched_filter(img.astype(np.float32).astype(np.float64), kern, sigma=10.0)
    assert r3.peaks == o.peaks
    fit = pkg.fit_all_stars(img - 5.0, psf, r3, 5, n_passes=2, fixed=dict(fwhm=2.5, bkg=0.0), ctx=ctx)
    order = [r3.peaks.index((int(a), int(b))) for a, b in zip(ys, xs)]
    np.testing.assert_allclose(fit.flux[order], fl, rtol=0.02)
    np.testing.assert_allclose(fit.y[order], ys, atol=0.05)
