"""Host-side contract of the Python mirror (argument errors, generator), CPU only."""
import numpy as np
import pytest

from _support import load_pkg

pkg = load_pkg()
sim = pkg.simulation


def test_simultaneous_argument_contract():
    psf = pkg.CircularGaussianPSF(fwhm=2.0)
    img = np.zeros((64, 64))
    cat = {"y": [32.0], "x": [32.0], "flux": [500.0]}
    f = pkg.fit_all_stars_simultaneous
    with pytest.raises(pkg.ArgumentError, match="only \\(y, x, flux\\)"):
        f(img, psf, cat, 5, model_rad=5)          # bkg and fwhm left free
    with pytest.raises(pkg.ArgumentError, match="LevenbergDamping"):
        f(img, psf, cat, 5, fixed={"fwhm": 2.0, "bkg": 0.0}, damping="levenberg", model_rad=5)
    with pytest.raises(pkg.ArgumentError, match="solver"):
        f(img, psf, cat, 5, fixed={"fwhm": 2.0, "bkg": 0.0}, solver="foo", model_rad=5)
    with pytest.raises(pkg.ArgumentError, match="requires inv_var"):
        f(img, psf, cat, 3, fixed={"fwhm": 2.0, "bkg": 0.0})
    with pytest.raises(pkg.ArgumentError, match="all model parameters are fixed"):
        f(img, psf, cat, 3, fixed={"y": 1, "x": 1, "flux": 1, "fwhm": 2.0, "bkg": 0.0}, model_rad=3)
    for kw in [dict(max_iter=0), dict(inner_iterations=0), dict(linear_tol=0.0), dict(max_trials=0),
               dict(model_rad_nsigma=0.0), dict(model_rad_max=1.0), dict(model_rad=-1)]:
        with pytest.raises(pkg.ArgumentError):
            f(img, psf, cat, 5, fixed={"fwhm": 2.0, "bkg": 0.0}, **{"model_rad": 5, **kw})
    with pytest.raises(pkg.ArgumentError, match="same size"):
        f(img, psf, cat, 5, fixed={"fwhm": 2.0, "bkg": 0.0}, inv_var=np.ones((3, 3)), model_rad=5)


def test_multipass_argument_contract():
    """fit_all_stars' argument errors are raised on the host, before any
    device work (psf_photometry_single.jl:279,305)."""
    psf = pkg.CircularGaussianPSF(fwhm=2.0)
    img = np.zeros((64, 64))
    cat = {"y": [32.0], "x": [32.0], "flux": [500.0]}
    f = pkg.fit_all_stars
    with pytest.raises(pkg.ArgumentError, match="n_passes"):
        f(img, psf, cat, 5, n_passes=0)
    with pytest.raises(pkg.ArgumentError, match="all model parameters are fixed"):
        f(img, psf, cat, 5, fixed={"y": 1, "x": 1, "flux": 1, "fwhm": 2.0, "bkg": 0.0})
    with pytest.raises(pkg.ArgumentError, match="same size"):
        f(img, psf, cat, 5, inv_var=np.ones((3, 3)))
    with pytest.raises(pkg.ArgumentError, match="no parameter"):
        f(img, psf, cat, 5, fixed={"alpha": 1.0})
    with pytest.raises(pkg.ArgumentError, match="unsupported PSF model"):
        f(img, object(), cat, 5)


def test_generator_is_seeded_and_follows_distributions():
    y, x, fl = sim.simulate_sources(7, 20000, (8, 248), (8, 248), (1e2, 1e5), "loguniform")
    y2, x2, fl2 = sim.simulate_sources(7, 20000, (8, 248), (8, 248), (1e2, 1e5), "loguniform")
    assert np.array_equal(y, y2) and np.array_equal(fl, fl2)
    assert y.min() >= 8 and y.max() <= 248
    lf = np.log10(fl)
    assert abs(lf.mean() - 3.5) < 0.03 and lf.min() >= 2 and lf.max() <= 5
    # power law dN/dF ~ F^-1.5 on [1e2, 1e6]: CDF closed form at the median
    _, _, pf = sim.simulate_sources(3, 20000, (0, 1), (0, 1), (1e2, 1e6), "powerlaw", 1.5)
    q = 1 - 1.5
    med = ((1e2 ** q + 0.5 * (1e6 ** q - 1e2 ** q))) ** (1 / q)
    assert abs(np.median(pf) / med - 1) < 0.05


def test_noise_moments_and_determinism():
    img = np.full((200, 300), 50.0)
    a = sim.add_noise(img, 11, 1.0, 4.0)
    b = sim.add_noise(img, 11, 1.0, 4.0)
    assert np.array_equal(a, b)
    assert abs(a.mean() - 50) < 0.1 and abs(a.var() - (50 + 16)) < 1.5
    small = sim.add_noise(np.full((100, 100), 3.0), 2, 1.0, 0.0)
    assert np.all(small == np.round(small)) and abs(small.mean() - 3) < 0.05 and abs(small.var() - 3) < 0.15


def test_render_matches_closed_form_gaussian():
    psf = pkg.CircularGaussianPSF(fwhm=2.5)
    img = sim.render_sources(psf, (40, 50), [20.3], [25.7], [1000.0], 7.0)
    gy, gx = np.meshgrid(np.arange(1, 41), np.arange(1, 51), indexing="ij")
    g = 1000.0 * 4 * np.log(2) / (np.pi * 6.25) * np.exp(-4 * np.log(2) * ((gy - 20.3) ** 2 + (gx - 25.7) ** 2) / 6.25)
    inside = (gy >= 13) & (gy <= 28) & (gx >= 18) & (gx <= 33)
    np.testing.assert_allclose(img[inside], g[inside], rtol=1e-12)
    assert np.all(img[~inside] == 0)


def test_epsf_grid_normalisation():
    e = pkg.moffat_epsf(fwhm=2.2, beta=2.5, size=101, oversampling=4, origin=(51.0, 51.0))
    assert e.data.shape == (101, 101)
    assert abs(e.data.sum() - 16.0) < 1e-9
    k = pkg.pixel_response_kernel(4)
    assert k.shape == (5, 5) and abs(k.sum() - 1) < 1e-15 and abs(k[0, 0] - 1 / 64) < 1e-15


def test_source_catalogue_forms():
    """_extract_source_catalog (psf_photometry_single.jl:152-197): a table, a
    per-star record list (measure_star_shapes: centroid.y, centroid.x,
    matched_filter_flux) and a MatchedFilterResult give the same initial
    (y, x, flux); a record list missing a field is an ArgumentError."""
    from collections import namedtuple
    cat = pkg.photometry._catalog
    y, x, f = np.array([10.5, 20.25]), np.array([11.0, 30.5]), np.array([100.0, 250.0])
    Cen = namedtuple("Cen", "y x")
    Rec = namedtuple("Rec", "centroid matched_filter_flux")
    forms = [
        {"y": y, "x": x, "flux": f},
        [Rec(Cen(a, b), c) for a, b, c in zip(y, x, f)],
        [{"centroid": {"y": a, "x": b}, "matched_filter_flux": c} for a, b, c in zip(y, x, f)],
        type("MF", (), {"peaks": np.column_stack([y, x]), "peak_fluxes": f})(),
    ]
    for s in forms:
        gy, gx, gf = cat(s)
        np.testing.assert_array_equal(gy, y)
        np.testing.assert_array_equal(gx, x)
        np.testing.assert_array_equal(gf, f)
    gy, gx, gf = cat({"y": y, "x": x})
    np.testing.assert_array_equal(gf, np.ones(2))
    with pytest.raises(pkg.ArgumentError, match="matched_filter_flux"):
        cat([{"centroid": {"y": 1.0, "x": 2.0}}])
    # a table that is itself a named tuple is a table, not a record list
    T = namedtuple("T", "y x flux")
    gy, _, _ = cat(T(y, x, f))
    np.testing.assert_array_equal(gy, y)


def test_solver_keyword():
    psf = pkg.CircularGaussianPSF(fwhm=2.0)
    call = pkg.photometry._SimCall
    kw = dict(fixed={"fwhm": 2.0, "bkg": 0.0}, model_rad=5)
    cat = {"y": [32.0], "x": [32.0], "flux": [500.0]}
    img = np.zeros((64, 64))
    assert call(img, psf, cat, 5, solver="lsqr", **kw).opts.solver == pkg._abi.CP_SOLVER_LSQR
    assert call(img, psf, cat, 5, solver="lsmr", **kw).opts.solver == pkg._abi.CP_SOLVER_LSMR
    assert call(img, psf, cat, 5, solver="cg", **kw).opts.solver == pkg._abi.CP_SOLVER_CGLS
    c = call(img, psf, cat, 5, show_trace=True, max_iter=7, max_trials=3, **kw)
    assert c.result.trace_cap == 21
