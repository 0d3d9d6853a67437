"""Oracle pinned against the reference's known answers and closed forms (CPU)."""
import math

import numpy as np
import pytest

from _support import f64, golden, load_pkg, oracle, ptr

pkg = load_pkg()
abi = pkg._abi
G = golden()


def eval_fg(model, py, px):
    lib = oracle()
    d = model.desc()
    prm = f64(model.params())
    n = abi.NPARAMS[model.kind]
    f = np.zeros(1)
    g = np.zeros(n)
    assert lib.or_evaluate_fg(d, ptr(prm, abi.f64p), py, px, ptr(f, abi.f64p), ptr(g, abi.f64p)) == 0
    return f[0], g


def evaluate(model, py, px):
    return oracle().or_evaluate(model.desc(), ptr(f64(model.params()), abi.f64p), py, px)


def test_circular_gaussian_known_answer():
    g = G["circular_gaussian"]
    m = pkg.CircularGaussianPSF(**g["params"])
    f, grad = eval_fg(m, *g["at"])
    assert f == pytest.approx(g["f"], rel=1e-14)
    assert evaluate(m, *g["at"]) == pytest.approx(g["f"], rel=1e-14)
    np.testing.assert_allclose(grad, g["G"], rtol=1e-12)


def test_circular_moffat_known_answer():
    g = G["circular_moffat"]
    m = pkg.CircularMoffatPSF(**g["params"])
    f, grad = eval_fg(m, *g["at"])
    assert f == pytest.approx(g["f"], rel=1e-14)
    np.testing.assert_allclose(grad, g["G"], rtol=1e-12)
    # fwhm <-> alpha relation (parametric_models.jl:689-692)
    assert pkg.CircularMoffatPSF.alpha_from_fwhm(g["fwhm"], 3.0) == pytest.approx(5.0, rel=1e-14)


def _image_psf(**kw):
    g = G["image_psf"]
    i = np.arange(1, 8)[:, None]
    j = np.arange(1, 8)[None, :]
    data = np.exp(-((i - 4) ** 2 + (j - 4) ** 2) / 5)
    return pkg.ImagePSF(data, oversampling=g["oversampling"], normalize=g["normalize"], **kw)


def test_image_psf_known_answers():
    g = G["image_psf"]
    m = _image_psf(**g["params"])
    assert m.data.sum() == pytest.approx(g["sum"], rel=1e-14)
    ext = m.extent()
    np.testing.assert_allclose(ext, g["extent"], rtol=1e-14)
    assert evaluate(m, *g["far_value"]["at"]) == g["far_value"]["f"]
    fm = g["fill_model"]
    mf = _image_psf(fill_value=fm["fill_value"], **fm["params"])
    assert evaluate(mf, *fm["at"]) == fm["f"]


def test_image_psf_gradient_matches_finite_differences():
    g = G["image_psf"]
    fd_ = g["fd_gradient"]
    m = _image_psf(**g["params"])
    py, px = fd_["at"]
    f, grad = eval_fg(m, py, px)
    assert f == pytest.approx(evaluate(m, py, px), rel=1e-14)
    h = fd_["h"]
    fd = np.zeros(4)
    for k, name in enumerate(m.props):
        mp = m.setproperties(**{name: getattr(m, name) + h})
        mm = m.setproperties(**{name: getattr(m, name) - h})
        fd[k] = (evaluate(mp, py, px) - evaluate(mm, py, px)) / (2 * h)
    np.testing.assert_allclose(grad, fd, rtol=fd_["rtol"], atol=fd_["atol"])


def test_bicubic_closed_forms():
    lib = oracle()
    out = np.zeros(3)
    g = G["bicubic_out_of_bounds"]
    data = np.asfortranarray(np.random.default_rng(0).random((7, 7)))
    flat = f64(data.ravel(order="F"))
    lib.or_bicubic(ptr(flat, abi.f64p), 7, 7, g["at"][0], g["at"][1], g["fill_value"], ptr(out, abi.f64p))
    assert list(out) == g["expect"]
    # at an interior node: value = data and d/dx = (f3 - f1)/2 (empirical_models.jl:219-223)
    for (iy, ix) in [(3, 4), (4, 2), (5, 5)]:
        lib.or_bicubic(ptr(flat, abi.f64p), 7, 7, float(iy), float(ix), 0.0, ptr(out, abi.f64p))
        assert out[0] == pytest.approx(data[iy - 1, ix - 1], abs=1e-15)
        assert out[2] == pytest.approx((data[iy - 1, ix] - data[iy - 1, ix - 2]) / 2, abs=1e-15)
        assert out[1] == pytest.approx((data[iy, ix - 1] - data[iy - 2, ix - 1]) / 2, abs=1e-15)
    # Keys a=-1/2 reproduces a quadratic exactly in the interior, with exact derivatives
    yy, xx = np.meshgrid(np.arange(1, 10), np.arange(1, 10), indexing="ij")
    q = 0.3 * yy ** 2 - 0.2 * xx * yy + 0.7 * xx ** 2 + yy - 2 * xx + 5
    qf = f64(q.ravel(order="F"))
    for (y, x) in [(3.3, 4.7), (5.5, 5.5), (6.9, 3.1)]:
        lib.or_bicubic(ptr(qf, abi.f64p), 9, 9, y, x, 0.0, ptr(out, abi.f64p))
        assert out[0] == pytest.approx(0.3 * y * y - 0.2 * x * y + 0.7 * x * x + y - 2 * x + 5, rel=1e-12)
        assert out[1] == pytest.approx(0.6 * y - 0.2 * x + 1, rel=1e-12)
        assert out[2] == pytest.approx(-0.2 * y + 1.4 * x - 2, rel=1e-12)


def _four_node_model(**kw):
    g = G["gridded_bilinear"]
    rng = np.random.default_rng(3)
    base = np.exp(-((np.arange(1, 12)[:, None] - 6) ** 2 + (np.arange(1, 12)[None, :] - 6) ** 2) / 6)
    nodes = [pkg.ImagePSF(base * (1 + 0.3 * rng.random((11, 11))), normalize=True) for _ in range(4)]
    return pkg.GriddedPSFModel(nodes, g["node_grid_y"], g["node_grid_x"], **kw), nodes


def corners(model, Y, X):
    lib = oracle()
    idx = np.zeros(4, dtype=np.int32)
    w, dy, dx = np.zeros(4), np.zeros(4), np.zeros(4)
    lib.or_grid_corners_dw(model.desc(), Y, X, ptr(idx, abi.i32p), ptr(w, abi.f64p), ptr(dy, abi.f64p),
                           ptr(dx, abi.f64p))
    return idx, w, dy, dx


def test_gridded_manual_bilinear_blend():
    g = G["gridded_bilinear"]
    Y, X = g["manual_point"]["Y"], g["manual_point"]["X"]
    m, nodes = _four_node_model(y=Y, x=X, flux=100.0, bkg=1.0)
    w = [(10 - X) * (10 - Y) / 100, X * (10 - Y) / 100, (10 - X) * Y / 100, X * Y / 100]
    py, px = 3.0, 8.0
    manual = 100 * sum(wk * evaluate(n.setproperties(y=Y, x=X, flux=1.0, bkg=0.0), py, px)
                       for wk, n in zip(w, nodes)) + 1.0
    assert evaluate(m, py, px) == pytest.approx(manual, rel=1e-12)
    idx, ww, _, _ = corners(m, Y, X)
    np.testing.assert_allclose(ww, w, rtol=1e-14)
    for Yv in np.linspace(0, 10, 11):
        for Xv in np.linspace(0, 10, 11):
            _, ww, _, _ = corners(m, Yv, Xv)
            assert np.all(ww >= 0) and ww.sum() == pytest.approx(1.0, rel=1e-14)


def test_gridded_edge_clamp_and_single_node():
    g = G["gridded_bilinear"]
    m, nodes = _four_node_model()
    for Y, X, ci in g["edge_cases"]:
        mm = m.setproperties(y=Y, x=X, flux=100.0, bkg=1.0)
        rec = nodes[ci - 1].setproperties(y=Y, x=X, flux=1.0, bkg=0.0)
        for (py, px) in [(Y, X), (Y + 1.3, X - 0.7)]:
            assert evaluate(mm, py, px) == pytest.approx(100 * evaluate(rec, py, px) + 1.0, rel=1e-12)
    oa = g["one_axis_outside"]
    _, ww, _, _ = corners(m, oa["Y"], oa["X"])
    assert int(np.sum(ww != 0)) == oa["n_active"]
    m1 = pkg.GriddedPSFModel([nodes[0]], [5.0], [5.0], y=-30.0, x=200.0, flux=7.0, bkg=2.0)
    rec = nodes[0].setproperties(y=-30.0, x=200.0, flux=1.0, bkg=0.0)
    assert evaluate(m1, -30.0, 200.0) == pytest.approx(7 * evaluate(rec, -30.0, 200.0) + 2, rel=1e-14)


def test_gridded_gradient_finite_differences():
    m, _ = _four_node_model(y=3.7, x=6.2, flux=50.0, bkg=0.5)
    h = 1e-6
    for (py, px) in [(4.0, 6.0), (2.0, 8.0), (5.0, 3.0)]:
        f, grad = eval_fg(m, py, px)
        fd = []
        for name in m.props:
            mp = m.setproperties(**{name: getattr(m, name) + h})
            mm = m.setproperties(**{name: getattr(m, name) - h})
            fd.append((evaluate(mp, py, px) - evaluate(mm, py, px)) / (2 * h))
        np.testing.assert_allclose(grad, fd, rtol=1e-6, atol=1e-7)


def test_analytic_gradients_finite_differences():
    for m in [pkg.CircularGaussianPSF(y=1.2, x=-0.7, fwhm=2.6, flux=300, bkg=4),
              pkg.CircularMoffatPSF(y=1.2, x=-0.7, alpha=2.1, beta=2.7, flux=300, bkg=4)]:
        for (py, px) in [(2.0, 1.0), (0.0, -3.0)]:
            _, grad = eval_fg(m, py, px)
            h = 1e-6
            fd = []
            for name in m.props:
                mp = m.setproperties(**{name: getattr(m, name) + h})
                mm = m.setproperties(**{name: getattr(m, name) - h})
                fd.append((evaluate(mp, py, px) - evaluate(mm, py, px)) / (2 * h))
            np.testing.assert_allclose(grad, fd, rtol=1e-6, atol=1e-9)


def test_gaussian_cog_is_analytic_integral():
    lib = oracle()
    m = pkg.CircularGaussianPSF(fwhm=2.5)
    for r in [1.0, 2.0, 5.0]:
        sig = 2.5 / math.sqrt(8 * math.log(2))
        assert lib.or_curve_of_growth(m.desc(), r) == pytest.approx(1 - math.exp(-r * r / (2 * sig * sig)), rel=1e-14)


def test_extract_errors():
    g = G["extract_errors"]
    cov = f64(np.array(g["cov"]).ravel())
    fi = np.array([k - 1 for k in g["free_idx_1based"]], dtype=np.int32)
    out = np.full(g["n_params"], np.nan)
    oracle().or_extract_errors(ptr(cov, abi.f64p), 2, ptr(fi, abi.i32p), g["n_params"], ptr(out, abi.f64p))
    np.testing.assert_allclose(out, g["expect"], rtol=1e-15)
