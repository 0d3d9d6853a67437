"""GPU parity of sequential multi-pass photometry (fit_all_stars,
src/photometry/psf_photometry_single.jl:267-469): the level-scheduled HIP path
through the C ABI against the FP64 oracle's one-star-at-a-time loop on the
same seeded fields.  Bars: float64 device mode within summation-order
rounding (1e-6 relative; the sums run in another order, so a star whose
lm_irls stops one iteration apart may move by ~x_tol); mixed precision
(float32 model and gradient, float64 sums and residual) within 1e-4 relative
flux and 1e-3 px, or 0.02 of the statistical error, for as many stars as the
oracle keeps there under a float32-sized perturbation of its input."""
import numpy as np
import pytest

import oracle_fit
from _support import load_pkg

pkg = load_pkg()
sim = pkg.simulation

pytestmark = pytest.mark.gpu

G2 = pkg.CircularGaussianPSF(fwhm=2.0)
M25 = pkg.CircularMoffatPSF(alpha=pkg.CircularMoffatPSF.alpha_from_fwhm(2.4, 2.8), beta=2.8)


@pytest.fixture(scope="module")
def ctx():
    return pkg.Context(0)


def _crowded(psf, shape=(192, 160), n=260, seed=11, bkg=25.0, noise=True):
    """A field with blended groups: n stars uniform over the image (some near
    the edges), log-uniform fluxes, Poisson noise; the catalogue is the truth
    perturbed by 0.3 px and 20% in flux."""
    ny, nx = shape
    y, x, f = sim.simulate_sources(seed, n, (1.0, ny), (1.0, nx), flux=(300.0, 3e4))
    img = sim.render_sources(psf, shape, y, x, f, 15) + bkg
    if noise:
        img = sim.add_noise(img, seed + 1)
    rng = np.random.default_rng(seed)
    cat = dict(y=y + rng.normal(0, 0.3, n), x=x + rng.normal(0, 0.3, n), flux=f * rng.uniform(0.8, 1.2, n))
    return img.astype(np.float32).astype(np.float64), cat


def _agree(g, o, rtol, sig):
    """Per star: flux within rtol relative (or sig statistical errors), y and
    x within 10 rtol px (or sig errors)."""
    scale = np.maximum(np.abs(o.flux), 1.0)
    return (((np.abs(g.flux - o.flux) <= rtol * scale) | (np.abs(g.flux - o.flux) <= sig * o.flux_err))
            & ((np.abs(g.y - o.y) <= max(rtol, 1e-6) * 10) | (np.abs(g.y - o.y) <= sig * o.y_err))
            & ((np.abs(g.x - o.x) <= max(rtol, 1e-6) * 10) | (np.abs(g.x - o.x) <= sig * o.x_err)))


DIAG = ("qfit", "crowding", "spread_model", "qfit_expected", "qfit_z", "spread_model_err")


def _diag_close(g, o, k, both):
    a, b = getattr(g, k)[both], getattr(o, k)[both]
    fin = ~np.isnan(b)
    return np.abs(a[fin] - b[fin]) <= 1e-3 * np.maximum(np.abs(b[fin]), 1e-2)


def _compare(g, o, rtol, frac=1.0, check_diag=True, sig=0.0, diag_frac=None):
    # a star at a convergence boundary may flip its converged flag, hence validity
    assert np.mean(g.valid == o.valid) >= 0.99, np.argwhere(g.valid != o.valid).ravel()
    both = g.valid & o.valid
    assert both.sum() > 0.8 * len(g.valid)
    scale = np.maximum(np.abs(o.flux), 1.0)
    ok = _agree(g, o, rtol, sig)
    bad = np.argwhere(both & ~ok).ravel()
    assert ok[both].mean() >= frac, (ok[both].mean(), bad[:5], (g.flux - o.flux)[bad[:5]] / scale[bad[:5]],
                                     (g.y - o.y)[bad[:5]])
    if frac == 1.0:
        # errors of well-measured stars: a near-singular J'WJ (a blend whose
        # width and flux trade off) amplifies rounding without bound
        sane = both & (o.flux_err < 0.5 * np.abs(o.flux))
        np.testing.assert_allclose(g.flux_err[sane], o.flux_err[sane], rtol=max(rtol, 1e-6) * 100,
                                   atol=1e-3 * np.median(o.flux_err[sane]))
        np.testing.assert_allclose(g.bkg[both], o.bkg[both], rtol=rtol, atol=rtol * 10)
        assert np.mean(g.n_iter == o.n_iter) > 0.95
        assert np.mean(g.converged == o.converged) > 0.98
    if check_diag:
        for k in DIAG:
            np.testing.assert_array_equal(np.isnan(getattr(g, k)[both]), np.isnan(getattr(o, k)[both]), err_msg=k)
            close = _diag_close(g, o, k, both)
            want = min(frac, 0.99) if diag_frac is None else diag_frac[k]
            assert close.size == 0 or close.mean() >= want, (k, close.mean(), want)
    assert g.n_passes == o.n_passes and g.n_failed == o.n_failed


@pytest.mark.parametrize("psf", [G2, M25], ids=["gauss", "moffat"])
@pytest.mark.parametrize("n_passes", [1, 3])
def test_crowded_fp64_matches_oracle(ctx, psf, n_passes):
    img, cat = _crowded(psf)
    iv = (1.0 / np.maximum(img, 1.0)).astype(np.float32).astype(np.float64)   # what the library receives
    kw = dict(n_passes=n_passes, inv_var=iv, max_iter=100)
    o = oracle_fit.fit_all(psf, img, cat, 5, **kw)
    g = pkg.fit_all_stars(img, psf, cat, 5, precision="fp64", ctx=ctx, **kw)
    print(f"valid {g.valid.sum()}/{len(g.valid)} n_iter {g.n_iter.sum()} vs {o.n_iter.sum()}", flush=True)
    _compare(g, o, 1e-6)
    np.testing.assert_allclose(g.residual, o.residual, rtol=1e-5, atol=1e-3)


@pytest.mark.parametrize("psf", [G2, M25], ids=["gauss", "moffat"])
def test_crowded_mixed_matches_oracle(ctx, psf):
    img, cat = _crowded(psf, seed=13)
    # the usual photometry set-up: known PSF shape and sky, free y, x, flux
    shape = {k: getattr(psf, k) for k in psf.props if k not in ("y", "x", "flux", "bkg")}
    kw = dict(n_passes=2, fixed=dict(bkg=25.0, **shape), max_iter=100)
    o = oracle_fit.fit_all(psf, img, cat, 5, **kw)
    g = pkg.fit_all_stars(img, psf, cat, 5, ctx=ctx, **kw)
    # The sequential fit of a blended field is sensitive to rounding: the
    # oracle itself, run on the image perturbed by float32-sized (1e-7
    # relative) noise, moves a percent or two of the stars by more than 1e-4.
    # The mixed path must agree with the oracle at least about as well as
    # the oracle agrees with itself under that perturbation.
    rng = np.random.default_rng(0)
    o3 = oracle_fit.fit_all(psf, img * (1 + 1e-7 * rng.standard_normal(img.shape)), cat, 5, **kw)
    both = o.valid & o3.valid
    self_frac = _agree(o3, o, 1e-4, 0.02)[both].mean()
    diag_frac = {k: (_diag_close(o3, o, k, both).mean() if _diag_close(o3, o, k, both).size else 1.0) - 0.02
                 for k in DIAG}
    print(f"oracle self-agreement under 1e-7 perturbation: {self_frac:.4f} {diag_frac}", flush=True)
    _compare(g, o, 1e-4, frac=min(0.995, self_frac - 0.01), sig=0.02, diag_frac=diag_frac)


def test_reference_integration_cases(ctx):
    """psf_photometry_single.jl:105-266 on the GPU: recovery, fixed
    parameters, the edge star, the rejected negative flux."""
    y = np.array([30.2, 70.7, 101.4, 45.9, 88.3])
    x = np.array([20.6, 90.1, 40.3, 111.8, 60.5])
    f = np.array([650.0, 820.0, 700.0, 880.0, 610.0])
    img = sim.render_sources(G2, (128, 128), y, x, f, 30) + 20.0
    r = pkg.fit_all_stars(img, G2, dict(y=y + 0.2, x=x - 0.2, flux=f * 0.9), 5, n_passes=3, fixed=dict(bkg=20.0),
                          max_iter=100, ctx=ctx)
    assert r.n_passes == 3 and r.n_failed == 0 and r.valid.all() and r.converged.all()
    np.testing.assert_allclose(r.flux, f, rtol=1e-3)
    np.testing.assert_allclose(r.y, y, atol=1e-3)
    assert np.all(np.isnan(r.qfit_expected)) and np.all(np.isfinite(r.crowding))
    np.testing.assert_array_equal(r.bkg_err, 0.0)
    r = pkg.fit_all_stars(img, G2, dict(y=y, x=x, flux=f), 5, n_passes=1, fixed=dict(x=32.0, y=32.0),
                          max_iter=100, ctx=ctx)
    np.testing.assert_array_equal(r.x, 32.0)
    np.testing.assert_array_equal(r.y_err, 0.0)
    edge = sim.render_sources(G2, (64, 64), [2.2, 32.3], [32.4, 31.8], [550.0, 520.0], 5) + 20.0
    r = pkg.fit_all_stars(edge, G2, dict(y=[1.0, 32.0], x=[32.0, 32.0], flux=[500.0, 500.0]), 5, n_passes=1,
                          max_iter=50, ctx=ctx)
    assert r.valid[1]
    # a negative catalogue flux over empty sky: whether the between-pass gate
    # drops it depends on where its fit lands (here a ~0 positive flux, kept
    # by the oracle too)
    bad = dict(y=np.r_[y, 50.0], x=np.r_[x, 50.0], flux=np.r_[f, -100.0])
    r = pkg.fit_all_stars(img, G2, bad, 5, n_passes=2, max_iter=100, ctx=ctx, precision="fp64")
    o = oracle_fit.fit_all(G2, img, bad, 5, n_passes=2, max_iter=100)
    np.testing.assert_array_equal(r.valid, o.valid)
    np.testing.assert_allclose(r.flux[:5], o.flux[:5], rtol=1e-7)   # the library takes the image as float32


def test_spread_model_cases(ctx):
    """psf_photometry_single.jl:269-321 on the GPU, against the oracle."""
    psf = pkg.CircularGaussianPSF(fwhm=2.5)
    fx = dict(bkg=20.0, fwhm=2.5)
    cat = dict(y=[30.4], x=[29.7], flux=[5000.0])
    sm = []
    for w in (2.5, 3.0, 4.0, 6.0):
        img = sim.render_sources(pkg.CircularGaussianPSF(fwhm=w), (60, 60), [30.4], [29.7], [5000.0], 1000) + 20.0
        iv = np.full(img.shape, 1 / 20.0)
        g = pkg.fit_all_stars(img, psf, cat, 8, n_passes=2, max_iter=100, fixed=fx, inv_var=iv, ctx=ctx,
                              precision="fp64")
        o = oracle_fit.fit_all(psf, img, cat, 8, n_passes=2, max_iter=100, fixed=fx, inv_var=iv)
        assert g.spread_model[0] == pytest.approx(o.spread_model[0], abs=1e-7)
        assert g.spread_model_err[0] == pytest.approx(o.spread_model_err[0], rel=1e-6)
        sm.append(g.spread_model[0])
    assert abs(sm[0]) < 3e-3 and all(s > 5e-4 for s in sm[1:]) and sm == sorted(sm)
    spike = np.full((40, 40), 20.0)
    spike[19, 19] += 3000.0
    r = pkg.fit_all_stars(spike, psf, dict(y=[20.0], x=[20.0], flux=[3000.0]), 8, n_passes=2, max_iter=100,
                          fixed=fx, inv_var=np.full(spike.shape, 1 / 20.0), ctx=ctx)
    assert r.valid[0] and r.spread_model[0] < -1e-3


def test_failures_empty_and_out_of_image(ctx):
    """Too few valid inv_var pixels -> invalid + n_failed (like the oracle);
    a star whose box misses the image -> invalid without a failure; an empty
    catalogue -> n_passes 0 and a 0 x 0 residual."""
    img = sim.render_sources(G2, (64, 64), [20.3, 44.1], [20.8, 43.2], [700.0, 650.0], 8) + 20.0
    iv = np.ones_like(img)
    iv[38:50, 37:49] = 0.0   # star 2's box (rows 39-50, columns 38-49) but one pixel
    iv[38, 37] = 1.0
    cat = dict(y=[20.3, 44.1, -30.0], x=[20.8, 43.2, 10.0], flux=[700.0, 650.0, 500.0])
    for kw in (dict(n_passes=1), dict(n_passes=2)):
        g = pkg.fit_all_stars(img, G2, cat, 5, inv_var=iv, max_iter=100, precision="fp64", ctx=ctx, **kw)
        o = oracle_fit.fit_all(G2, img, cat, 5, inv_var=iv, max_iter=100, **kw)
        assert g.n_failed == o.n_failed == 1
        np.testing.assert_array_equal(g.valid, [True, False, False])
        np.testing.assert_array_equal(g.valid, o.valid)
        np.testing.assert_allclose(g.residual, o.residual, rtol=1e-6, atol=1e-4)
    r = pkg.fit_all_stars(img, G2, dict(y=np.zeros(0), x=np.zeros(0)), 5, ctx=ctx)
    assert r.n_passes == 0 and r.residual.shape == (0, 0)
    with pytest.raises(pkg.ArgumentError, match="position"):
        pkg.fit_all_stars(img, G2, dict(y=[np.nan], x=[3.0]), 5, ctx=ctx)


def _epsf(n_nodes):
    e = pkg.moffat_epsf(fwhm=2.2, beta=2.5, size=41, oversampling=4)
    if n_nodes == 1:
        return pkg.GriddedPSFModel([e], [0.0], [0.0])
    e2 = pkg.moffat_epsf(fwhm=2.7, beta=3.0, size=41, oversampling=4)
    return pkg.GriddedPSFModel([e, e2, e2, e], [1.0, 1.0, 192.0, 192.0], [1.0, 160.0, 1.0, 160.0])


@pytest.mark.parametrize("n_nodes", [1, 4])
def test_crowded_epsf_matches_oracle(ctx, n_nodes):
    """ImagePSF / GriddedPSFModel stars (y, x, flux, bkg free), fp64 against
    the oracle to summation-order rounding, and the mixed path against the
    oracle's own self-agreement."""
    psf = _epsf(n_nodes)
    img, cat = _crowded(psf, seed=17)
    iv = (1.0 / np.maximum(img, 1.0)).astype(np.float32).astype(np.float64)
    kw = dict(n_passes=2, inv_var=iv, max_iter=100)
    o = oracle_fit.fit_all(psf, img, cat, 4, **kw)
    g = pkg.fit_all_stars(img, psf, cat, 4, precision="fp64", ctx=ctx, **kw)
    print(f"epsf nodes {n_nodes}: valid {g.valid.sum()}/{len(g.valid)} n_iter {g.n_iter.sum()} vs "
          f"{o.n_iter.sum()}", flush=True)
    _compare(g, o, 1e-6)
    np.testing.assert_allclose(g.residual, o.residual, rtol=1e-5, atol=1e-3)
    kw = dict(n_passes=2, fixed=dict(bkg=25.0), max_iter=100)
    o = oracle_fit.fit_all(psf, img, cat, 4, **kw)
    g = pkg.fit_all_stars(img, psf, cat, 4, ctx=ctx, **kw)
    rng = np.random.default_rng(0)
    o3 = oracle_fit.fit_all(psf, img * (1 + 1e-7 * rng.standard_normal(img.shape)), cat, 4, **kw)
    both = o.valid & o3.valid
    self_frac = _agree(o3, o, 1e-4, 0.02)[both].mean()
    diag_frac = {k: (_diag_close(o3, o, k, both).mean() if _diag_close(o3, o, k, both).size else 1.0) - 0.02
                 for k in DIAG}
    _compare(g, o, 1e-4, frac=min(0.995, self_frac - 0.01), sig=0.02, diag_frac=diag_frac)


def _gridded_gauss():
    g = np.arange(-5, 6, dtype=float)
    node = np.exp(pkg.psf.GAUSS_PRE * (g[:, None] ** 2 + g[None, :] ** 2) / 4.0)
    return pkg.GriddedPSFModel([pkg.ImagePSF(node, oversampling=1, normalize=True)], [0.0], [0.0])


@pytest.mark.parametrize("kind", ["gauss", "gridded"])
def test_isolated_stars_sequential_equals_simultaneous(ctx, kind):
    """test/photometry/psf_photometry_simultaneous.jl:219-253 on the GPU, both
    paths in float64."""
    psf, fixed = (G2, dict(fwhm=2.0, bkg=0.0)) if kind == "gauss" else (_gridded_gauss(), dict(bkg=0.0))
    ys = np.array([20.3, 40.7, 64.2, 90.6, 110.1])
    xs = np.array([30.4, 95.2, 60.8, 20.5, 100.9])
    fl = np.array([650.0, 720.0, 880.0, 610.0, 790.0])
    img = sim.render_sources(psf, (128, 128), ys, xs, fl, 5)
    cat = dict(y=ys, x=xs, flux=np.full(5, 400.0))
    seq = pkg.fit_all_stars(img, psf, cat, 5, n_passes=1, fixed=fixed, max_iter=200, precision="fp64", ctx=ctx)
    simr = pkg.fit_all_stars_simultaneous(img, psf, cat, 5, fixed=fixed, max_iter=40, inner_iterations=10,
                                          x_tol=1e-8, model_rad=5, precision="fp64", ctx=ctx)
    assert simr.valid.all() and seq.valid.all()
    np.testing.assert_allclose(simr.flux, seq.flux, rtol=1e-7)    # the image travels as float32
    np.testing.assert_allclose(simr.y, seq.y, rtol=0, atol=1e-7)
    np.testing.assert_allclose(seq.flux, fl, rtol=1e-6)


@pytest.mark.parametrize("n_nodes", [1, 4])
def test_fit_star_batch_epsf_matches_oracle(ctx, n_nodes):
    """fit_star over ePSF cutouts (psf_fitting.jl:1086-1210): each cutout
    holds one star of the model at its own position in the frame of the node
    grid; (y, x, flux, bkg) free; both precisions against the oracle."""
    import ctypes as C
    from _support import f64, oracle, ptr
    psf = _epsf(n_nodes)
    rng = np.random.default_rng(5)
    n = 40
    cuts, th0 = [], []
    for i in range(n):
        s = int(rng.choice([11, 13, 15]))
        y, x = (s + 1) / 2 + rng.uniform(-0.5, 0.5), (s + 1) / 2 + rng.uniform(-0.5, 0.5)
        f, b = rng.uniform(500, 5000), rng.uniform(5, 50)
        c = sim.render_sources(psf, (s, s), [y], [x], [f], 40) + b
        cuts.append(sim.add_noise(c, 100 + i).astype(np.float32).astype(np.float64))
        th0.append([y + rng.normal(0, 0.2), x + rng.normal(0, 0.2), f * rng.uniform(0.8, 1.2), b])
    th0 = np.array(th0)
    abi = pkg._abi
    ny = np.array([c.shape[0] for c in cuts], np.int32)
    off = np.zeros(n, np.int64)
    off[1:] = np.cumsum(ny.astype(np.int64) ** 2)[:-1]
    pix = np.concatenate([c.T.ravel() for c in cuts]).astype(np.float32)
    tho = np.zeros((n, 4))
    sto = (abi.LMStat * n)()
    o = abi.lm_opts_default()
    oracle().or_fit_star_batch(psf.desc(), ptr(pix, abi.f32p), ptr(off, abi.i64p), ptr(ny, abi.i32p),
                               ptr(ny, abi.i32p), None, ptr(f64(th0), abi.f64p), 0b1111, n, C.byref(o),
                               ptr(tho, abi.f64p), None, sto, 0)
    th, cov, st = pkg.fit_star_batch(psf, cuts, th0, ctx=ctx)
    assert all(s.status == 0 for s in st) and all(s.status == 0 for s in sto)
    rel = np.abs(th[:, 2] - tho[:, 2]) / tho[:, 2]
    dp = np.maximum(np.abs(th[:, 0] - tho[:, 0]), np.abs(th[:, 1] - tho[:, 1]))
    assert np.mean((rel <= 1e-4) & (dp <= 1e-3)) >= 0.95, (rel.max(), dp.max())
    assert np.median(rel) < 1e-6
