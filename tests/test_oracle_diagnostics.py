"""Post-fit per-star diagnostics (qfit, qfit_expected, qfit_z, crowding,
spread_model, spread_model_err) of the oracle and the library's host-side
spread_model reference disk, pinned on the CPU against closed forms and the
reference's own test invariants
(test/photometry/psf_photometry_simultaneous.jl:336-383,
test/photometry/psf_photometry_single.jl:269-321)."""
import math

import numpy as np
import pytest
import scipy.special as sp

import fixtures
import oracle_fit
from _support import f64, load_pkg, oracle, ptr

pkg = load_pkg()
abi = pkg._abi

# Hogg & Lang six-Gaussian approximation of a unit-integral exponential disk,
# variances in units of the half-light radius squared (the table the reference
# quotes at src/photometry/psf_photometry_diagnostics.jl:14-15).
AMP = np.array([1.99485977e-4, 2.61612679e-3, 1.89726655e-2, 1.00186544e-1, 3.68534484e-1, 5.09490694e-1])
VAR = np.array([1.20078965e-3, 8.84526493e-3, 3.91463084e-2, 1.39976817e-1, 4.60962500e-1, 1.50159566e0])
B_EXP = 1.67834699   # half-light radius of exp(-r/h) in units of h


def test_mixture_table_is_an_exponential_disk():
    """The table's enclosed-flux curve follows 1 - (1 + r/h) exp(-r/h) (to the
    1.5% the six-term fit reaches) with half-light radius 1: a swapped
    amplitude/variance or a wrong entry fails."""
    assert abs(AMP.sum() - 1) < 1e-6
    for r in np.linspace(0.05, 4.0, 80):
        mix = np.sum(AMP * (1 - np.exp(-r * r / (2 * VAR))))
        x = r * B_EXP
        exact = 1 - (1 + x) * math.exp(-x)
        assert abs(mix - exact) < 0.015, (r, mix, exact)
    assert abs(np.sum(AMP * (1 - np.exp(-1 / (2 * VAR)))) - 0.5) < 0.01


def band_profile(m, a):
    """int_{-1/2}^{1/2} exp(-a k^2) cos(2 pi k m) dk by completing the square:
    exp(-pi^2 m^2 / a) sqrt(pi / a) Re erf(sqrt(a)/2 + i pi m / sqrt(a))."""
    if math.pi ** 2 * m * m / a > 690:
        return 0.0
    sa = math.sqrt(a)
    return (math.exp(-math.pi ** 2 * m * m / a) * math.sqrt(math.pi / a)
            * sp.erf(sa / 2 + 1j * math.pi * m / sa).real)


def disk_kernel(fwhm, half):
    re = fwhm / 16 * B_EXP
    n = 2 * half + 1
    K = np.zeros((n, n))
    for a_c, v in zip(AMP, VAR):
        a = 2 * math.pi ** 2 * re * re * v
        iv = np.array([band_profile(m, a) for m in range(-half, half + 1)])
        K += a_c * np.outer(iv, iv)
    return K / K.sum()


def test_band_profile_closed_form_against_direct_quadrature():
    """Above pi^2 m^2 / a = 690 the reference drops the term (its erf form
    overflows there); the true integral is then at most the sinc ringing of a
    nearly flat spectrum, O(a / m^2)."""
    from scipy.integrate import quad
    for a in (0.05, 0.7, 3.0, 20.0):
        for m in (0, 1, 2, 5):
            want = quad(lambda k: math.exp(-a * k * k) * math.cos(2 * math.pi * k * m), -0.5, 0.5,
                        limit=200, epsabs=1e-14)[0]
            if math.pi ** 2 * m * m / a > 690:
                assert band_profile(m, a) == 0.0 and abs(want) < a / (m * m)
            else:
                assert abs(band_profile(m, a) - want) < 1e-12


@pytest.mark.parametrize("fwhm,half", [(1.5, 3), (2.5, 5), (4.0, 8), (6.0, 4), (2.2, 12)])
def test_spread_kernel_oracle_and_library_match_closed_form(fwhm, half):
    want = disk_kernel(fwhm, half)
    n = 2 * half + 1
    ko = np.zeros(n * n)
    assert oracle().or_spread_kernel(fwhm, half, ptr(ko, abi.f64p)) == 0
    lib = abi.load_library()
    kl = np.zeros(n * n)
    assert lib.cp_spread_kernel(fwhm, half, ptr(kl, abi.f64p)) == 0
    np.testing.assert_allclose(ko.reshape(n, n, order="F"), want, rtol=0, atol=1e-13)
    np.testing.assert_allclose(kl.reshape(n, n, order="F"), want, rtol=0, atol=1e-13)


def test_spread_kernel_invariants():
    """sum 1, four-fold symmetric, peak at the centre, and a broader disk
    takes more from the centre (psf_photometry_single.jl:323-334)."""
    lib = abi.load_library()
    cen = {}
    for fw in (1.5, 2.5, 4.0, 6.0, 2.0):
        n = 11
        k = np.zeros(n * n)
        assert lib.cp_spread_kernel(fw, 5, ptr(k, abi.f64p)) == 0
        K = k.reshape(n, n, order="F")
        assert abs(K.sum() - 1) < 1e-13
        np.testing.assert_allclose(K, K[::-1, :], atol=1e-15)
        np.testing.assert_allclose(K, K[:, ::-1], atol=1e-15)
        np.testing.assert_allclose(K, K.T, atol=1e-15)
        assert np.unravel_index(np.argmax(K), K.shape) == (5, 5)
        cen[fw] = K[5, 5]
    assert cen[6.0] < cen[2.0]
    bad = np.zeros(1)
    assert lib.cp_spread_kernel(-1.0, 0, ptr(bad, abi.f64p)) != 0
    assert lib.cp_spread_kernel(2.0, -1, ptr(bad, abi.f64p)) != 0


# --------------------------------------------------------------- whole fit
def _box(y, x, r, shape):
    ny, nx = shape
    y0, y1 = max(math.floor(y - r), 1), min(math.ceil(y + r), ny)
    x0, x1 = max(math.floor(x - r), 1), min(math.ceil(x + r), nx)
    return slice(y0 - 1, y1), slice(x0 - 1, x1)


def test_qfit_is_global_residual_based():
    """qfit = sum |global residual| over the clamped fit box / flux
    (psf_photometry_simultaneous.jl test :336-354)."""
    rng = np.random.default_rng(5)
    shape = (64, 64)
    ys, xs = fixtures.separated_positions(rng, 3, shape, 8, 6)
    fs = rng.uniform(500, 800, 3)
    img = fixtures.render_gauss(shape, ys, xs, fs, 2.0, 5)
    psf = pkg.CircularGaussianPSF(fwhm=2.0)
    r = oracle_fit.fit(psf, img, ys, xs, fs.copy(), 5, max_iter=40, model_rad=5.0)
    assert r.valid.all()
    for i in range(3):
        by, bx = _box(r.y[i], r.x[i], 5, shape)
        want = np.abs(r.residual[by, bx]).sum() / r.flux[i]
        assert r.qfit[i] == pytest.approx(want, rel=1e-6, abs=1e-12)
    # no inv_var: qfit_expected, qfit_z and spread_model_err stay NaN
    assert np.isnan(r.qfit_expected).all() and np.isnan(r.qfit_z).all()
    assert np.isnan(r.spread_model_err).all()
    assert np.isfinite(r.spread_model).all()


def test_qfit_expected_and_z_closed_form():
    """With a constant inverse variance w over a box of n pixels:
    qfit_expected = sqrt(2/pi) n w^-1/2 / flux and
    qfit_z = (qfit flux - sqrt(2/pi) d n s) / (sqrt((1 - 2/pi) n) s d),
    s = w^-1/2, d = sqrt(1 - p/n) (p = 3 free parameters)."""
    rng = np.random.default_rng(9)
    shape = (80, 80)
    ys, xs = fixtures.separated_positions(rng, 4, shape, 12, 14)
    fs = rng.uniform(2000, 4000, 4)
    img = fixtures.render_gauss(shape, ys, xs, fs, 2.5, 8) + rng.normal(0, 3.0, shape)
    w = 1 / 9.0
    iv = np.full(shape, w)
    psf = pkg.CircularGaussianPSF(fwhm=2.5)
    r = oracle_fit.fit(psf, img, ys, xs, fs.copy(), 6, inv_var=iv, max_iter=40, model_rad=8.0)
    assert r.valid.all()
    s = 1 / math.sqrt(w)
    c = math.sqrt(2 / math.pi)
    for i in range(4):
        by, bx = _box(r.y[i], r.x[i], 6, shape)
        n = (by.stop - by.start) * (bx.stop - bx.start)
        assert r.qfit_expected[i] == pytest.approx(c * n * s / r.flux[i], rel=1e-12)
        d = math.sqrt(1 - 3 / n)
        q = r.qfit[i] * r.flux[i]
        want = (q - c * d * n * s) / (math.sqrt((1 - 2 / math.pi) * n) * s * d)
        assert r.qfit_z[i] == pytest.approx(want, rel=1e-9, abs=1e-9)
        # pure noise residuals: |z| is O(1)
        assert abs(r.qfit_z[i]) < 5
        assert np.isfinite(r.spread_model_err[i]) and r.spread_model_err[i] > 0
    # isolated stars: the dirty and clean fluxes agree, so crowding ~ 0
    assert np.all(np.abs(r.crowding) < 0.05)


def _spread_field():
    img = np.zeros((90, 90))
    pts = [(20.5, 20.5), (20.5, 68.5), (68.5, 20.5)]
    for y, x in pts:
        img += fixtures.render_gauss(img.shape, [y], [x], [4000.0], 2.5, 20)
    img += fixtures.render_gauss(img.shape, [68.5], [68.5], [4000.0], 5.0, 25)
    ys = np.array([p[0] for p in pts] + [68.5])
    xs = np.array([p[1] for p in pts] + [68.5])
    return img, ys, xs


def test_spread_model_point_sources_zero_extended_positive():
    """(psf_photometry_simultaneous.jl test :356-383): three isolated point
    sources fitted with the matching PSF give spread_model ~ 0, the injected
    broadened source a clearly positive value; a broader explicit reference
    disk gives the extended source a larger value."""
    img, ys, xs = _spread_field()
    iv = np.full(img.shape, 1 / 20.0)
    psf = pkg.CircularGaussianPSF(fwhm=2.5)
    r = oracle_fit.fit(psf, img, ys, xs, np.full(4, 4000.0), 7, inv_var=iv, max_iter=40)
    assert r.valid.all()
    assert np.all(np.abs(r.spread_model[:3]) < 3e-3)
    assert r.spread_model[3] > 5e-3
    assert np.all(np.isfinite(r.spread_model_err)) and np.all(r.spread_model_err > 0)
    big = oracle_fit.fit(psf, img, ys, xs, np.full(4, 4000.0), 7, inv_var=iv, max_iter=40,
                         spread_model_fwhm=8.0)
    assert big.spread_model[3] > r.spread_model[3]
    # the default reference width is the effective-area FWHM of the PSF: for a
    # circular Gaussian rendered on pixels it is close to the PSF's own FWHM
    near = oracle_fit.fit(psf, img, ys, xs, np.full(4, 4000.0), 7, inv_var=iv, max_iter=40,
                          spread_model_fwhm=2.5)
    np.testing.assert_allclose(near.spread_model, r.spread_model, rtol=0.05, atol=2e-4)


def test_spread_model_sharper_than_psf_is_negative():
    """A single hot pixel fitted with a broad PSF: spread_model < 0
    (psf_photometry_single.jl test :302-308)."""
    img = np.zeros((40, 40))
    img[19, 19] = 3000.0
    iv = np.full(img.shape, 1 / 20.0)
    psf = pkg.CircularGaussianPSF(fwhm=2.5)
    r = oracle_fit.fit(psf, img, [20.0], [20.0], [3000.0], 8, inv_var=iv, max_iter=100)
    assert r.valid[0]
    assert r.spread_model[0] < -1e-3


def test_crowding_grows_with_a_fitted_neighbour():
    """crowding = 2.5 log10(dirty / clean): the clean flux has every other
    star's fitted model removed, the dirty flux does not, so a fitted bright
    neighbour makes it positive, and an unfitted one (left in the residual)
    changes neither."""
    shape = (60, 60)
    base = fixtures.render_gauss(shape, [30.2], [30.6], [3000.0], 2.5, 12)
    nb = fixtures.render_gauss(shape, [30.2], [35.1], [3000.0], 2.5, 12)
    psf = pkg.CircularGaussianPSF(fwhm=2.5)
    kw = dict(max_iter=40, model_rad=12.0)
    a = oracle_fit.fit(psf, base, [30.2], [30.6], [3000.0], 6, **kw)
    b = oracle_fit.fit(psf, base + nb, [30.2, 30.2], [30.6, 35.1], [3000.0, 3000.0], 6, **kw)
    c = oracle_fit.fit(psf, base + nb, [30.2], [30.6], [3000.0], 6, **kw)
    assert abs(a.crowding[0]) < 1e-6
    assert b.crowding[0] > 0.01
    assert abs(c.crowding[0]) < 1e-6


def test_diagnostics_nan_for_invalid_and_inactive():
    img, ys, xs = _spread_field()
    psf = pkg.CircularGaussianPSF(fwhm=2.5)
    ys2 = np.concatenate([ys, [-50.0]])
    xs2 = np.concatenate([xs, [-50.0]])
    r = oracle_fit.fit(psf, img, ys2, xs2, np.full(5, 4000.0), 7, max_iter=20, model_rad=20.0)
    assert not r.valid[4]
    for k in ("qfit", "qfit_expected", "qfit_z", "crowding", "spread_model", "spread_model_err"):
        assert np.isnan(getattr(r, k)[4]), k
