"""Oracle fit_all_stars (sequential multi-pass photometry,
src/photometry/psf_photometry_single.jl:267-469), pinned on the CPU:

* the reference's own integration tests (test/photometry/psf_photometry_single.jl:105-321);
* noiseless single stars recover the truth that generated them;
* the orchestration (flux order, +-fit_rad boxes, add-back on later passes,
  subtraction, error extraction, between-pass validation) against the same
  steps composed by hand from the oracle's fit_star and evaluate, which the
  model and LM tests pin separately;
* the residual as image minus the fitted models.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_fit
from _support import f64, load_pkg, oracle, ptr

pkg = load_pkg()
abi = pkg._abi
sim = pkg.simulation

G2 = pkg.CircularGaussianPSF(fwhm=2.0)


def _field(shape, n, min_sep, border, flux, seed, psf=G2, bkg=20.0, model_radius=30):
    """Noiseless image of n stars with pairwise separation >= min_sep, at
    least `border` pixels from the edges (simulate_image's min_separation /
    border), background bkg."""
    rng = np.random.default_rng(seed)
    ny, nx = shape
    ys, xs = [], []
    while len(ys) < n:
        y, x = rng.uniform(border, ny - border), rng.uniform(border, nx - border)
        if all((y - a) ** 2 + (x - b) ** 2 >= min_sep ** 2 for a, b in zip(ys, xs)):
            ys.append(y)
            xs.append(x)
    ys, xs = np.array(ys), np.array(xs)
    fs = rng.uniform(flux[0], flux[1], n)
    img = sim.render_sources(psf, shape, ys, xs, fs, model_radius) + bkg
    return img, dict(y=ys, x=xs, flux=fs)


def _box(y, x, r, ny, nx):
    return (max(math.floor(y - r), 1), min(math.ceil(y + r), ny), max(math.floor(x - r), 1),
            min(math.ceil(x + r), nx))


def _render(psf, prm, box):
    ylo, yhi, xlo, xhi = box
    lib = oracle()
    p = f64(prm)
    d = psf.desc()
    return np.array([[lib.or_evaluate(d, ptr(p, abi.f64p), float(gy), float(gx)) for gx in range(xlo, xhi + 1)]
                     for gy in range(ylo, yhi + 1)])


def _fit_star(psf, cut, theta, mask, oy, ox, **lm):
    """Oracle fit_star on a box whose pixel (1, 1) is image pixel (oy + 1, ox + 1)."""
    lib = oracle()
    th = f64(theta).copy()
    th[0] -= oy
    th[1] -= ox
    o = abi.lm_opts_default()
    for k, v in lm.items():
        setattr(o, k, v)
    ny, nx = cut.shape
    pix = f64(cut.ravel(order="F"))
    p = bin(mask).count("1")
    out = np.zeros(len(th))
    cov = np.zeros(p * p)
    st = abi.LMStat()
    rc = lib.or_fit_star(psf.desc(), ptr(pix, abi.f64p), ny, nx, None, ptr(th, abi.f64p), mask, C.byref(o),
                         ptr(out, abi.f64p), ptr(cov, abi.f64p), C.byref(st))
    assert rc == 0
    out[0] += oy
    out[1] += ox
    return out, cov.reshape(p, p), st


# --------------------------------------------------------------------------
# the reference's integration tests (psf_photometry_single.jl:105-266)
# --------------------------------------------------------------------------
def test_noiseless_recovery():
    img, src = _field((128, 128), 5, 7, 8, (600.0, 900.0), seed=42)
    r = oracle_fit.fit_all(G2, img, src, 5, n_passes=3, fixed=dict(bkg=20.0), max_iter=100)
    assert r.n_passes == 3 and r.n_failed == 0
    assert r.valid.sum() == 5 and r.converged.all()
    np.testing.assert_allclose(r.flux, src["flux"], rtol=0.10)
    np.testing.assert_allclose(r.y, src["y"], atol=0.01)
    np.testing.assert_allclose(r.x, src["x"], atol=0.01)
    assert np.all(np.isfinite(r.qfit)) and np.all(r.qfit >= 0)
    assert np.all(np.isnan(r.qfit_expected)) and np.all(np.isnan(r.qfit_z))
    assert np.all(np.isfinite(r.crowding)) and np.all(r.crowding >= 0)
    np.testing.assert_array_equal(r.bkg, 20.0)
    np.testing.assert_array_equal(r.bkg_err, 0.0)


def test_single_pass_with_inv_var():
    img, src = _field((128, 128), 5, 10, 10, (600.0, 900.0), seed=43, model_radius=5)
    r = oracle_fit.fit_all(G2, img, src, 5, n_passes=1, inv_var=np.ones_like(img), max_iter=100)
    assert r.n_passes == 1 and r.valid.all() and r.converged.all()
    assert np.all(np.isfinite(r.qfit_expected)) and np.all(r.qfit_expected > 0)
    assert np.all(np.isfinite(r.qfit_z))


def test_fixed_background_and_positions():
    img, src = _field((64, 64), 5, 10, 8, (500.0, 800.0), seed=44, model_radius=5)
    r = oracle_fit.fit_all(G2, img, src, 5, n_passes=1, fixed=dict(bkg=20.0), max_iter=100)
    np.testing.assert_allclose(r.bkg, 20.0)
    np.testing.assert_array_equal(r.bkg_err, 0.0)
    np.testing.assert_allclose(r.flux, src["flux"], rtol=0.02)
    r = oracle_fit.fit_all(G2, img, src, 5, n_passes=1, fixed=dict(x=32.0, y=32.0), max_iter=100)
    np.testing.assert_array_equal(r.x, 32.0)
    np.testing.assert_array_equal(r.y, 32.0)
    np.testing.assert_array_equal(r.x_err, 0.0)
    np.testing.assert_array_equal(r.y_err, 0.0)


def test_edge_star_and_multipass():
    img = sim.render_sources(G2, (64, 64), [2.2, 32.3], [32.4, 31.8], [550.0, 520.0], 5) + 20.0
    r = oracle_fit.fit_all(G2, img, dict(y=[1.0, 32.0], x=[32.0, 32.0], flux=[500.0, 500.0]), 5, n_passes=1,
                           max_iter=50)
    assert r.valid[1]
    img, src = _field((64, 64), 5, 5, 8, (400.0, 700.0), seed=46, model_radius=5)
    r = oracle_fit.fit_all(G2, img, src, 5, n_passes=3, max_iter=100)
    assert r.n_passes == 3 and r.valid.sum() == 5 and r.converged.all()


def test_catalogue_fluxes_and_invalid_flux_rejected():
    img, src = _field((64, 64), 3, 15, 10, (500.0, 700.0), seed=47, model_radius=5)
    r = oracle_fit.fit_all(G2, img, dict(y=src["y"], x=src["x"], flux=np.full(3, 400.0)), 5, n_passes=1,
                           max_iter=200)
    assert r.converged.all()
    np.testing.assert_allclose(r.flux, src["flux"], rtol=0.05)
    img, src = _field((64, 64), 5, 10, 8, (400.0, 700.0), seed=48, model_radius=5)
    bad = dict(y=np.r_[src["y"], 50.0], x=np.r_[src["x"], 50.0], flux=np.r_[src["flux"], -100.0])
    r = oracle_fit.fit_all(G2, img, bad, 5, n_passes=2, max_iter=100)
    assert r.valid.sum() == 5 and not r.valid[5]


def _point(fwhm=2.5, y=30.4, x=29.7, shape=(60, 60), flux=5000.0):
    return sim.render_sources(pkg.CircularGaussianPSF(fwhm=fwhm), shape, [y], [x], [flux], 1000) + 20.0


def test_spread_model():
    """psf_photometry_single.jl:269-321."""
    psf = pkg.CircularGaussianPSF(fwhm=2.5)
    fx = dict(bkg=20.0, fwhm=2.5)
    cat = dict(y=[30.4], x=[29.7], flux=[5000.0])
    img = _point()
    iv = np.full(img.shape, 1 / 20.0)
    r = oracle_fit.fit_all(psf, img, cat, 8, n_passes=2, max_iter=100, fixed=fx, inv_var=iv)
    assert r.valid[0] and abs(r.spread_model[0]) < 3e-3
    assert np.isfinite(r.spread_model_err[0]) and r.spread_model_err[0] > 0
    r = oracle_fit.fit_all(psf, img, cat, 8, n_passes=2, max_iter=100, fixed=fx)
    assert np.isfinite(r.spread_model[0]) and np.isnan(r.spread_model_err[0])
    sm = [oracle_fit.fit_all(psf, _point(fwhm=w), cat, 8, n_passes=2, max_iter=100, fixed=fx,
                             inv_var=iv).spread_model[0] for w in (2.5, 3.0, 4.0, 6.0)]
    assert all(s > 5e-4 for s in sm[1:]) and sm == sorted(sm)
    spike = np.full((40, 40), 20.0)
    spike[19, 19] += 3000.0
    r = oracle_fit.fit_all(psf, spike, dict(y=[20.0], x=[20.0], flux=[3000.0]), 8, n_passes=2, max_iter=100,
                           fixed=fx, inv_var=np.full(spike.shape, 1 / 20.0))
    assert r.valid[0] and r.spread_model[0] < -1e-3
    ext = _point(fwhm=4.0)
    a = oracle_fit.fit_all(psf, ext, cat, 8, n_passes=2, max_iter=100, fixed=fx, inv_var=iv)
    b = oracle_fit.fit_all(psf, ext, cat, 8, n_passes=2, max_iter=100, fixed=fx, inv_var=iv, spread_model_fwhm=8.0)
    assert a.spread_model[0] > 0 and b.spread_model[0] > a.spread_model[0]


# --------------------------------------------------------------------------
# pins beyond the reference's tests
# --------------------------------------------------------------------------
@pytest.mark.parametrize("psf", [G2, pkg.CircularMoffatPSF(alpha=pkg.CircularMoffatPSF.alpha_from_fwhm(2.5, 3.0),
                                                            beta=3.0)])
def test_noiseless_isolated_star_recovers_truth(psf):
    """One star rendered with the same model over a box wider than the fit
    box: one pass returns the generating parameters and leaves a zero
    residual over the fit box.  A second pass over a zero background returns
    them again (the models it adds back and subtracts carry the background
    term, and the box follows the fitted centre, so a nonzero background
    would shift it: the reference does the same)."""
    img = sim.render_sources(psf, (50, 60), [25.37], [31.81], [4000.0], 1000)
    tight = dict(max_iter=200, x_tol=1e-12, f_tol=1e-12, g_tol=1e-12)
    cat = dict(y=[25.0], x=[31.5], flux=[3000.0])
    r = oracle_fit.fit_all(psf, img + 7.0, cat, 6, n_passes=1, **tight)
    assert r.valid[0] and r.converged[0]
    assert r.y[0] == pytest.approx(25.37, abs=1e-7) and r.x[0] == pytest.approx(31.81, abs=1e-7)
    assert r.flux[0] == pytest.approx(4000.0, rel=1e-7) and r.bkg[0] == pytest.approx(7.0, abs=1e-6)
    ylo, yhi, xlo, xhi = _box(25.0, 31.5, 6, 50, 60)
    np.testing.assert_allclose(r.residual[ylo - 1:yhi, xlo - 1:xhi], 0.0, atol=1e-4)
    # the second pass's box takes one more row, never subtracted in the first:
    # negligible for the Gaussian, a 1e-5 pixel and 3e-4 flux pull from the
    # Moffat's tail
    tol = 1e-7 if isinstance(psf, pkg.CircularGaussianPSF) else 1e-3
    r = oracle_fit.fit_all(psf, img, cat, 6, n_passes=2, fixed=dict(bkg=0.0), **tight)
    assert r.y[0] == pytest.approx(25.37, abs=tol) and r.flux[0] == pytest.approx(4000.0, rel=tol)
    assert r.n_iter[0] > 1


def _by_hand(psf, img, src, fit_rad, n_passes, mask, by_flux=True, **lm):
    """The sequential algorithm composed from oracle fit_star / evaluate
    (by_flux=False: catalogue order instead of brightest first)."""
    names = psf.props
    n = len(src["y"])
    prm = np.tile(psf.params(), (n, 1))
    for k, nme in enumerate(names):
        if nme in src:
            prm[:, k] = src[nme]
    rf = names.index("flux")
    R = np.array(img, np.float64)
    ny, nx = R.shape
    valid = np.ones(n, bool)
    conv = np.zeros(n, bool)
    err = np.where([(mask >> k) & 1 for k in range(len(names))], np.nan, 0.0) * np.ones((n, 1))
    free = [k for k in range(len(names)) if (mask >> k) & 1]
    nit = np.zeros(n, int)
    for p in range(1, n_passes + 1):
        order = sorted(range(n), key=lambda i: -prm[i, rf]) if by_flux else range(n)
        for i in order:
            if not valid[i]:
                continue
            b = _box(prm[i, 0], prm[i, 1], fit_rad, ny, nx)
            ylo, yhi, xlo, xhi = b
            if p > 1:
                R[ylo - 1:yhi, xlo - 1:xhi] += _render(psf, prm[i], b)
            th, cov, st = _fit_star(psf, R[ylo - 1:yhi, xlo - 1:xhi].copy(), prm[i], mask, ylo - 1, xlo - 1, **lm)
            prm[i] = th
            err[i, free] = np.sqrt(np.maximum(0, np.diag(cov)))
            conv[i] = st.converged
            nit[i] += st.iterations
            R[ylo - 1:yhi, xlo - 1:xhi] -= _render(psf, th, b)
        valid &= conv & np.isfinite(prm[:, 0]) & np.isfinite(prm[:, 1]) & (prm[:, rf] > 0)
    return prm, err, valid, nit, R


@pytest.mark.parametrize("n_passes", [1, 2])
def test_orchestration_matches_composed_steps(n_passes):
    """A blended group (boxes overlapping, fluxes spread so the order matters)
    plus an isolated star: every output equals the hand-composed sequence of
    fit_star / add_star! / subtract_star! calls to rounding (the coordinate
    shift of the cutout fit costs a few ulps)."""
    src = dict(y=np.array([20.3, 22.1, 19.0, 40.2]), x=np.array([20.6, 23.0, 24.4, 41.7]),
               flux=np.array([900.0, 2500.0, 1400.0, 1000.0]))
    img = sim.render_sources(G2, (60, 60), src["y"], src["x"], src["flux"], 30) + 20.0
    guess = dict(y=src["y"] + [0.3, -0.2, 0.25, 0.1], x=src["x"] + [-0.2, 0.3, 0.1, -0.3],
                 flux=src["flux"] * [0.8, 1.1, 1.2, 0.9], bkg=np.full(4, 20.0))
    mask = 0b01011   # y, x, flux free; fwhm and bkg held at the catalogue's values
    r = oracle_fit.fit_all(G2, img, guess, 4, n_passes=n_passes, fixed=dict(fwhm=2.0, bkg=20.0), max_iter=100)
    prm, err, valid, nit, R = _by_hand(G2, img, guess, 4, n_passes, mask, max_iter=100)
    np.testing.assert_allclose(r.y, prm[:, 0], rtol=0, atol=1e-9)
    np.testing.assert_allclose(r.x, prm[:, 1], rtol=0, atol=1e-9)
    np.testing.assert_allclose(r.flux, prm[:, 3], rtol=1e-10)
    np.testing.assert_allclose(r.flux_err, err[:, 3], rtol=1e-7)
    np.testing.assert_allclose(r.y_err, err[:, 0], rtol=1e-7)
    np.testing.assert_array_equal(r.valid, valid)
    np.testing.assert_array_equal(r.n_iter, nit)
    np.testing.assert_allclose(r.residual, R.astype(np.float32), rtol=1e-6, atol=1e-4)
    # stars 1-3 overlap, so the order matters: catalogue order gives other numbers
    alt = _by_hand(G2, img, guess, 4, n_passes, mask, by_flux=False, max_iter=100)[0]
    assert np.max(np.abs(alt[:3, 3] - r.flux[:3])) > 1.0
    np.testing.assert_allclose(alt[3], prm[3], rtol=1e-12)   # the isolated star does not care


def test_residual_is_image_minus_fitted_models():
    img, src = _field((96, 96), 6, 16, 10, (500.0, 900.0), seed=50, model_radius=5)
    r = oracle_fit.fit_all(G2, img, src, 5, n_passes=1, fixed=dict(bkg=20.0), max_iter=100)
    want = np.array(img, np.float64)
    for i in range(6):
        prm = [r.y[i], r.x[i], 2.0, r.flux[i], 20.0]
        b = _box(src["y"][i], src["x"][i], 5, 96, 96)
        want[b[0] - 1:b[1], b[2] - 1:b[3]] -= _render(G2, prm, b)
    np.testing.assert_allclose(r.residual, want, atol=2e-4)


def test_failed_fit_is_undone_and_counted():
    """A star whose box has too few valid inv_var pixels throws in fit_star:
    it is marked invalid, counted in n_failed and keeps its catalogue values
    (errors NaN), while its neighbour fits as before."""
    img, src = _field((64, 64), 2, 25, 12, (600.0, 800.0), seed=51, model_radius=8)
    iv = np.ones_like(img)
    r1 = oracle_fit.fit_all(G2, img, src, 5, n_passes=1, inv_var=iv, max_iter=100)
    assert r1.n_failed == 0 and r1.valid.all()
    iv2 = iv.copy()
    b = _box(src["y"][1], src["x"][1], 5, 64, 64)
    iv2[b[0] - 1:b[1], b[2] - 1:b[3]] = 0.0
    iv2[b[0] - 1, b[2] - 1] = 1.0   # one valid pixel < 5 free parameters
    r = oracle_fit.fit_all(G2, img, src, 5, n_passes=1, inv_var=iv2, max_iter=100)
    assert r.n_failed == 1 and not r.valid[1] and r.valid[0]
    assert np.isnan(r.flux_err[1]) and r.flux[1] == src["flux"][1]


def test_empty_catalogue_and_argument_errors():
    img = np.full((20, 20), 1.0)
    r = oracle_fit.fit_all(G2, img, dict(y=np.zeros(0), x=np.zeros(0)), 3, n_passes=2)
    assert r.n_passes == 0 and r.n_failed == 0
    with pytest.raises(ValueError, match="n_passes"):
        oracle_fit.fit_all(G2, img, dict(y=[5.0], x=[5.0]), 3, n_passes=0)
    with pytest.raises(ValueError, match="fixed"):
        oracle_fit.fit_all(G2, img, dict(y=[5.0], x=[5.0]), 3, fixed=dict(y=1, x=1, fwhm=2, flux=1, bkg=0))
    with pytest.raises(ValueError, match="position"):
        oracle_fit.fit_all(G2, img, dict(y=[np.nan], x=[5.0]), 3)


# --------------------------------------------------------------------------
# the sequential path against the simultaneous one
# (test/photometry/psf_photometry_simultaneous.jl:219-253, 356-383)
# --------------------------------------------------------------------------
def _gridded_gauss():
    """The reference's single-node ePSF of CircularGaussianPSF(fwhm=2):
    render(psf) over its extent (+-5 px), unit oversampling, normalised."""
    g = np.arange(-5, 6, dtype=float)
    node = np.exp(pkg.psf.GAUSS_PRE * (g[:, None] ** 2 + g[None, :] ** 2) / 4.0)
    return pkg.GriddedPSFModel([pkg.ImagePSF(node, oversampling=1, normalize=True)], [0.0], [0.0])


@pytest.mark.parametrize("kind", ["gauss", "gridded"])
def test_isolated_stars_sequential_equals_simultaneous(kind):
    psf, fixed = (G2, dict(fwhm=2.0, bkg=0.0)) if kind == "gauss" else (_gridded_gauss(), dict(bkg=0.0))
    img, src = _field((128, 128), 5, 15, 10, (600.0, 900.0), seed=42, psf=psf, bkg=20.0, model_radius=5)
    sub = img - 20.0
    cat = dict(y=src["y"], x=src["x"], flux=np.full(5, 400.0))
    seq = oracle_fit.fit_all(psf, sub, cat, 5, n_passes=1, fixed=fixed, max_iter=200)
    sim_ = oracle_fit.fit(psf, sub, cat["y"], cat["x"], cat["flux"], 5, max_iter=40, inner_iterations=10, x_tol=1e-8,
                          model_rad=5.0, fixed_bkg=0.0)
    assert sim_.valid.all() and seq.valid.all()
    np.testing.assert_allclose(sim_.flux, seq.flux, rtol=1e-10)
    np.testing.assert_allclose(sim_.y, seq.y, rtol=0, atol=1e-10)
    np.testing.assert_allclose(sim_.x, seq.x, rtol=0, atol=1e-10)
    np.testing.assert_allclose(seq.flux, src["flux"], rtol=1e-10)


def test_spread_model_sequential_matches_simultaneous():
    psf = pkg.CircularGaussianPSF(fwhm=2.5)
    pts = [(20.5, 20.5), (20.5, 68.5), (68.5, 20.5)]
    ys = [p[0] for p in pts] + [68.5]
    xs = [p[1] for p in pts] + [68.5]
    img = sim.render_sources(psf, (90, 90), ys[:3], xs[:3], [4000.0] * 3, 1000)
    img += sim.render_sources(pkg.CircularGaussianPSF(fwhm=5.0), (90, 90), [68.5], [68.5], [4000.0], 1000)
    iv = np.full(img.shape, 1 / 20.0)
    cat = dict(y=ys, x=xs, flux=[4000.0] * 4)
    seq = oracle_fit.fit_all(psf, img, cat, 7, n_passes=3, fixed=dict(fwhm=2.5, bkg=0.0), max_iter=100, inv_var=iv)
    simr = oracle_fit.fit(psf, img, ys, xs, [4000.0] * 4, 7, inv_var=iv, max_iter=40, fixed_bkg=0.0, model_rad=0.0)
    for i in range(3):
        assert seq.valid[i] and simr.valid[i]
        assert abs(seq.spread_model[i]) < 3e-3 and abs(simr.spread_model[i]) < 3e-3
        assert abs(seq.spread_model[i] - simr.spread_model[i]) < 2e-3
    assert seq.spread_model[3] > 5e-3 and simr.spread_model[3] > 5e-3
    assert seq.spread_model[3] == pytest.approx(simr.spread_model[3], rel=0.25)
