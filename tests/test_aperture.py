"""Exact pixel / circle overlap and the pixel-integrated curve of growth that
model_rad = :auto uses for ImagePSF / GriddedPSFModel
(src/apertures.jl:52-166, src/curve_of_growth.jl:143-176,
_model_radii psf_photometry_simultaneous.jl:209-243), library and oracle,
pinned against areas and integrals that the geometry fixes (CPU)."""
import math

import numpy as np
import pytest
from scipy.integrate import quad

from _support import f64, load_pkg, oracle, ptr

pkg = load_pkg()
abi = pkg._abi


def _both():
    lib = abi.load_library()
    return [lib.cp_pixel_overlap, oracle().or_pixel_overlap]


def _chord_area(dx, dy, r):
    """Area of the unit pixel at (dx, dy) inside the circle, by adaptive
    quadrature of the covered vertical extent."""
    def ext(x):
        h = math.sqrt(max(0.0, r * r - x * x))
        return max(0.0, min(dy + 0.5, h) - max(dy - 0.5, -h))
    a, b = max(dx - 0.5, -r), min(dx + 0.5, r)
    if b <= a:
        return 0.0
    pts = [p for y in (dy - 0.5, dy + 0.5) if abs(y) < r for p in (-math.sqrt(r * r - y * y), math.sqrt(r * r - y * y))
           if a < p < b]
    return quad(ext, a, b, points=pts or None, epsabs=1e-13, epsrel=1e-13, limit=200)[0]


@pytest.mark.parametrize("r", [0.3, 0.71, 1.0, 2.5, 7.3, 40.0])
@pytest.mark.parametrize("cy,cx", [(0.0, 0.0), (0.37, -0.21), (0.5, 0.5)])
def test_overlaps_tile_the_circle(r, cy, cx):
    """Pixels tile the plane, so the overlaps of every pixel sum to pi r^2."""
    L = int(math.ceil(r)) + 2
    ii, jj = np.meshgrid(np.arange(-L, L + 1), np.arange(-L, L + 1), indexing="ij")
    for f in _both():
        s = sum(f(float(j) - cx, float(i) - cy, r) for i, j in zip(ii.ravel(), jj.ravel()))
        assert s == pytest.approx(math.pi * r * r, rel=1e-12, abs=1e-12)


def test_overlap_matches_quadrature_and_special_cases():
    rng = np.random.default_rng(3)
    for f in _both():
        for _ in range(300):
            r = rng.uniform(0.2, 12.0)
            ang = rng.uniform(0, 2 * math.pi)
            d = r + rng.uniform(-1.0, 1.0)
            dx, dy = d * math.cos(ang), d * math.sin(ang)
            assert f(dx, dy, r) == pytest.approx(_chord_area(dx, dy, r), abs=1e-11)
        assert f(0.1, -0.05, 0.3) == pytest.approx(math.pi * 0.09, rel=1e-13)   # circle inside the pixel
        assert f(1.0, 2.0, 5.0) == 1.0                                          # pixel inside the circle
        assert f(6.0, 0.0, 5.0) == 0.0
        # a pixel centred on the rim of a huge circle: half covered
        assert f(1.0e4, 0.0, 1.0e4) == pytest.approx(0.5, abs=1e-4)


def _epsf_models():
    e = pkg.moffat_epsf(fwhm=2.2, beta=2.5, size=61, oversampling=4)
    e2 = pkg.moffat_epsf(fwhm=2.8, beta=3.0, size=61, oversampling=4)
    one = pkg.GriddedPSFModel([e], [0.0], [0.0])
    four = pkg.GriddedPSFModel([e, e2, e2, e], [-50.0, -50.0, 60.0, 60.0], [-40.0, 70.0, -40.0, 70.0])
    return [e, one, four]


def test_library_curve_of_growth_matches_oracle():
    lib = abi.load_library()
    radii = f64(np.arange(1, 31))
    for m in _epsf_models():
        d = m.desc()
        got = np.zeros(30)
        assert lib.cp_curve_of_growth(d, 30, ptr(radii, abi.f64p), ptr(got, abi.f64p)) == abi.CP_OK
        want = np.array([oracle().or_curve_of_growth(d, float(r)) for r in radii])
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-15)
        assert np.all(np.diff(got) > -1e-12)
        # half-integer and fractional radii take the same path
        rr = f64([0.5, 1.5, 2.25, 7.0])
        g2 = np.zeros(4)
        assert lib.cp_curve_of_growth(d, 4, ptr(rr, abi.f64p), ptr(g2, abi.f64p)) == abi.CP_OK
        np.testing.assert_allclose(g2, [oracle().or_curve_of_growth(d, float(r)) for r in rr], rtol=1e-12)
    bad = f64([2.0, 1.0])
    out = np.zeros(2)
    assert lib.cp_curve_of_growth(_epsf_models()[0].desc(), 2, ptr(bad, abi.f64p), ptr(out, abi.f64p)) != abi.CP_OK


def test_curve_of_growth_of_a_flat_model_is_the_aperture_area():
    """A constant ImagePSF (value 1 over +-50 px): every aperture sum is the
    aperture's area pi r^2, which pins the bounding pixels and the weights
    apart from the model evaluation."""
    flat = pkg.ImagePSF(np.ones((101, 101)), oversampling=1)
    lib = abi.load_library()
    radii = f64([0.4, 1.0, 2.5, 3.0, 7.77, 20.0, 45.0])
    got = np.zeros(len(radii))
    assert lib.cp_curve_of_growth(flat.desc(), len(radii), ptr(radii, abi.f64p), ptr(got, abi.f64p)) == abi.CP_OK
    np.testing.assert_allclose(got, np.pi * radii ** 2, rtol=1e-12)
    for r in radii:
        assert oracle().or_curve_of_growth(flat.desc(), float(r)) == pytest.approx(math.pi * r * r, rel=1e-12)


def test_pixel_curve_of_growth_of_analytic_models():
    """The generic aperture sum of a Gaussian / Moffat approaches the
    closed-form curve of growth: the model is point-sampled at pixel centres,
    which biases small apertures low by a few percent (partial pixels take
    their centre value), and reaches the closed form as r grows."""
    lib = oracle()
    alpha = pkg.CircularMoffatPSF.alpha_from_fwhm(4.0, 3.0)
    for model in (pkg.CircularGaussianPSF(fwhm=4.0), pkg.CircularMoffatPSF(alpha=alpha, beta=3.0)):
        d = model.desc()
        for r, tol in ((1.0, 0.06), (2.0, 0.03), (3.5, 8e-3), (6.0, 2e-3), (10.0, 2e-4), (25.0, 2e-6)):
            a = lib.or_curve_of_growth(d, r)
            p = lib.or_curve_of_growth_pixels(d, r)
            assert p <= a and p == pytest.approx(a, rel=tol), (model, r)
    g = pkg.CircularGaussianPSF(fwhm=4.0).desc()
    assert lib.or_curve_of_growth_pixels(g, 15.0) == pytest.approx(1.0, abs=1e-10)


def test_epsf_curve_of_growth_follows_its_moffat():
    """The Moffat ePSF's curve of growth is that of the Moffat convolved with
    the pixel response; past a few FWHM it agrees with the closed form."""
    lib = abi.load_library()
    e = pkg.moffat_epsf(fwhm=2.2, beta=2.5, size=101, oversampling=4)
    m = pkg.CircularMoffatPSF(alpha=pkg.CircularMoffatPSF.alpha_from_fwhm(2.2, 2.5), beta=2.5)
    radii = f64([4.0, 6.0, 9.0, 12.0])
    got = np.zeros(4)
    assert lib.cp_curve_of_growth(e.desc(), 4, ptr(radii, abi.f64p), ptr(got, abi.f64p)) == abi.CP_OK
    want = [oracle().or_curve_of_growth(m.desc(), float(r)) for r in radii]
    np.testing.assert_allclose(got, want, rtol=2e-2)


def test_oracle_auto_radii_for_epsf():
    """model_rad = :auto with an ePSF: radii lie in [R_fit, R_cap], grow with
    the catalogue flux and with a lower noise floor."""
    lib = oracle()
    e = pkg.GriddedPSFModel([pkg.moffat_epsf(fwhm=2.2, beta=2.5, size=101, oversampling=4)], [0.0], [0.0])
    fl = f64([10.0, 100.0, 1e3, 1e4, 1e5, 1e6])
    out = {}
    for sig in (1.0, 10.0):
        w = f64(np.full(4096, 1.0 / sig ** 2))
        o = np.zeros(6, np.int32)
        assert lib.or_model_radii(e.desc(), 0.0, 1.0, 3, 30, ptr(w, abi.f64p), 4096, ptr(fl, abi.f64p), 6,
                                  ptr(o, abi.i32p)) == 0
        assert o.min() >= 3 and o.max() <= 30 and np.all(np.diff(o) >= 0)
        out[sig] = o
    assert np.all(out[1.0] >= out[10.0]) and out[1.0][-1] > out[1.0][0]
