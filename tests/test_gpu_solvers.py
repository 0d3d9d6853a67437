"""The simultaneous fit's inner solvers on the GPU: LSMR (the reference's
solver = :lsmr, psf_photometry_simultaneous.jl:992,1164) and CG on the damped
normal equations (the build's :cg), each against the FP64 oracle running the
same solver, plus the reference's own solver-agreement tests
(test/photometry/psf_photometry_simultaneous.jl:219-292) run on the device."""
import numpy as np
import pytest

import fixtures
import oracle_fit
from _support import load_pkg

pkg = load_pkg()
abi = pkg._abi
sim = pkg.simulation

pytestmark = pytest.mark.gpu

CODE = {"lsqr": abi.CP_SOLVER_LSQR, "lsmr": abi.CP_SOLVER_LSMR, "cg": abi.CP_SOLVER_CGLS}


@pytest.fixture(scope="module")
def ctx():
    return pkg.Context(0)


def _agree(r, ro, flux_rtol, pos_atol):
    g = r.valid & ro.valid.astype(bool)
    rel = np.abs(r.flux[g] - ro.flux[g]) / np.abs(ro.flux[g])
    dp = np.maximum(np.abs(r.y[g] - ro.y[g]), np.abs(r.x[g] - ro.x[g]))
    return np.mean((rel <= flux_rtol) & (dp <= pos_atol)), rel, dp


@pytest.mark.parametrize("solver", ["lsmr", "cg"])
@pytest.mark.parametrize("case", ["gauss256", "epsf384"])
def test_solver_matches_oracle_with_same_solver(ctx, solver, case):
    print(f"case: {case} solver {solver}", flush=True)
    if case == "gauss256":
        cfg = sim.config_gauss_small(seed=1)
        cfg["image"] = cfg["image"].astype(np.float32).astype(np.float64)
    else:
        n = int(round(2e5 * (384 * 384) / 4096 ** 2))
        cfg = sim.config_epsf_crowded(seed=5, shape=(384, 384), n=n, max_iter=10)
    src = cfg["sources"]
    ro = oracle_fit.fit(cfg["psf"], np.asarray(cfg["image"], np.float64), src["y"], src["x"], src["flux"],
                        cfg["fit_rad"], inv_var=cfg["inv_var"], model_rad=cfg["model_rad"], max_iter=cfg["max_iter"],
                        fixed_bkg=0.0, solver=CODE[solver])
    kw = dict(fixed=cfg["fixed"], inv_var=cfg["inv_var"], model_rad=cfg["model_rad"], max_iter=cfg["max_iter"],
              solver=solver, ctx=ctx)
    r64 = pkg.fit_all_stars_simultaneous(cfg["image"], cfg["psf"], src, cfg["fit_rad"], precision="fp64", **kw)
    f64, rel, dp = _agree(r64, ro, 1e-6, 1e-6)
    print(f"  fp64: {f64:.5f} within 1e-6, max rel {rel.max():.2e}, max dpos {dp.max():.2e}, passes gpu "
          f"{r64.n_passes} oracle {ro.n_passes}, inner its gpu {r64.stats['n_inner']}", flush=True)
    assert np.array_equal(r64.valid, ro.valid.astype(bool)) and f64 == 1.0
    assert r64.n_passes == ro.n_passes
    r32 = pkg.fit_all_stars_simultaneous(cfg["image"], cfg["psf"], src, cfg["fit_rad"], precision="mixed", **kw)
    f32, rel, dp = _agree(r32, ro, 1e-4, 1e-3)
    print(f"  mixed: {f32:.5f} within 1e-4 / 1e-3 px, median rel {np.median(rel):.2e}", flush=True)
    assert f32 >= 0.99
    # a different solver gives a different (if close) answer
    rq = pkg.fit_all_stars_simultaneous(cfg["image"], cfg["psf"], src, cfg["fit_rad"], precision="fp64",
                                        **dict(kw, solver="lsqr"))
    assert not np.array_equal(rq.flux, r64.flux)


@pytest.mark.parametrize("kind", ["gauss", "gridded"])
def test_isolated_lsmr_and_cg_equal_lsqr(ctx, kind):
    """:219-253 -- on isolated noiseless stars LSMR (and CG) agree with LSQR to
    1e-9."""
    g = np.arange(-5, 6, dtype=float)
    if kind == "gauss":
        psf, fixed = pkg.CircularGaussianPSF(fwhm=2.0), {"fwhm": 2.0, "bkg": 0.0}
    else:
        node = np.exp(pkg.psf.GAUSS_PRE * (g[:, None] ** 2 + g[None, :] ** 2) / 4.0)
        psf = pkg.GriddedPSFModel([pkg.ImagePSF(node, oversampling=1, normalize=True)], [0.0], [0.0])
        fixed = {"bkg": 0.0}
    ys = np.array([20.3, 40.7, 64.2, 90.6, 110.1])
    xs = np.array([30.4, 95.2, 60.8, 20.5, 100.9])
    fl = np.array([650.0, 720.0, 880.0, 610.0, 790.0])
    img = sim.render_sources(psf, (128, 128), ys, xs, fl, 5)
    cat = dict(y=ys, x=xs, flux=np.full(5, 400.0))
    kw = dict(fixed=fixed, max_iter=40, inner_iterations=10, x_tol=1e-8, model_rad=5, precision="fp64", ctx=ctx)
    rq = pkg.fit_all_stars_simultaneous(img, psf, cat, 5, **kw)
    for solver in ("lsmr", "cg"):
        r = pkg.fit_all_stars_simultaneous(img, psf, cat, 5, solver=solver, **kw)
        assert r.valid.all()
        np.testing.assert_allclose(r.flux, rq.flux, rtol=1e-9)
        np.testing.assert_allclose(r.y, rq.y, rtol=0, atol=1e-9)
        np.testing.assert_allclose(r.x, rq.x, rtol=0, atol=1e-9)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_crowded_lsqr_lsmr_agree_with_mid_fit_freezing(ctx, precision):
    """:255-292 on the device: a crowded, blended Gaussian field where most
    stars freeze before max_iter; LSQR and LSMR agree in flux to 1e-2
    (norm-wise, as Julia's isapprox on vectors) and in position with median
    < 1e-5 px, 90th percentile < 1e-3 px and maximum < 2e-2 px."""
    img, ys, xs, fs = fixtures.crowded_field(11, min_sep=2.0)
    psf = pkg.CircularGaussianPSF(fwhm=2.0)
    cat = dict(y=ys, x=xs, flux=fs.copy())
    kw = dict(fixed={"fwhm": 2.0, "bkg": 0.0}, max_iter=25, inner_iterations=10, x_tol=1e-8, f_tol=0.0,
              model_rad=5, precision=precision, ctx=ctx)
    a = pkg.fit_all_stars_simultaneous(img, psf, cat, 5, **kw)
    b = pkg.fit_all_stars_simultaneous(img, psf, cat, 5, solver="lsmr", **kw)
    assert np.sum(a.n_iter[a.valid] < a.n_passes) > 0.5 * a.valid.sum()
    g = a.valid & b.valid
    assert g.sum() > 0.8 * len(g)
    assert np.linalg.norm(a.flux[g] - b.flux[g]) <= 1e-2 * max(np.linalg.norm(a.flux[g]), np.linalg.norm(b.flux[g]))
    for d in (np.abs(a.y[g] - b.y[g]), np.abs(a.x[g] - b.x[g])):
        assert np.median(d) < 1e-5 and np.quantile(d, 0.9) < 1e-3 and d.max() < 2e-2
