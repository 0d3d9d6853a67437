"""fit_star(model, image, inds; fixed, inv_var) on the GPU
(src/psf/psf_fitting.jl:229-308): a sub-region fit keeps the image's
coordinates, including for a multi-node GriddedPSFModel whose node blend
depends on the absolute star position.  Checked against the oracle fitting
the whole image with the pixels outside the region masked (inverse variance
0), which evaluates the model in image coordinates directly."""
import ctypes as C

import numpy as np
import pytest

from _support import f64, load_pkg, oracle, ptr

pkg = load_pkg()
abi = pkg._abi
sim = pkg.simulation

pytestmark = pytest.mark.gpu


def _four_nodes(side):
    e = pkg.moffat_epsf(fwhm=2.0, beta=2.5, size=41, oversampling=4)
    e2 = pkg.moffat_epsf(fwhm=3.2, beta=3.5, size=41, oversampling=4)
    return pkg.GriddedPSFModel([e, e2, e2, e], [1.0, 1.0, float(side), float(side)],
                               [1.0, float(side), 1.0, float(side)])


def test_fit_star_subregion_keeps_image_coordinates():
    ctx = pkg.Context(0)
    side = 96
    psf = _four_nodes(side)
    yt, xt, ft, bt = 61.3, 40.7, 4000.0, 12.0
    img = sim.render_sources(psf, (side, side), [yt], [xt], [ft], 40) + bt
    img = sim.add_noise(img, 17).astype(np.float32).astype(np.float64)
    iv = np.full((side, side), 1.0 / 16.0)
    rows, cols = slice(50, 73), slice(29, 52)          # 0-based: image rows 51..73, columns 30..52
    init = psf.setproperties(y=yt + 0.3, x=xt - 0.2, flux=ft * 0.9, bkg=bt * 1.1)
    best, res = pkg.fit_star(init, img, (rows, cols), inv_var=iv, ctx=ctx)
    assert res.converged
    # oracle: the whole image, weights zero outside the region
    ivm = np.zeros_like(iv)
    ivm[rows, cols] = iv[rows, cols]
    th = np.zeros(4)
    cov = np.zeros(16)
    st = abi.LMStat()
    o = abi.lm_opts_default()
    rc = oracle().or_fit_star(init.desc(), ptr(f64(img.T.ravel()), abi.f64p), side, side,
                              ptr(f64(ivm.T.ravel()), abi.f64p), ptr(f64(init.params()), abi.f64p), 0b1111,
                              C.byref(o), ptr(th, abi.f64p), ptr(cov, abi.f64p), C.byref(st))
    assert rc == 0
    assert best.flux == pytest.approx(th[2], rel=1e-4)
    assert best.y == pytest.approx(th[0], abs=1e-3) and best.x == pytest.approx(th[1], abs=1e-3)
    assert best.bkg == pytest.approx(th[3], rel=1e-3)
    # the blend really depends on where the star is: the same cutout fitted as
    # if it sat at the image origin gives a different flux
    cut = img[rows, cols]
    th_shift, _, _ = pkg.fit_star_batch(psf, [cut], np.array([[yt + 0.3 - 50, xt - 0.2 - 29, ft * 0.9, bt * 1.1]]),
                                        inv_var=[iv[rows, cols]], ctx=ctx)
    assert abs(th_shift[0, 2] - th[2]) > 1e-3 * th[2]


def test_show_trace_prints_the_reference_trial_lines(capsys):
    """show_trace (psf_photometry_simultaneous.jl:1238-1246): one line per
    damping trial with the live fraction, the cost change, lambda, the
    gradient cosine and the outcome, matching the fit's own trace."""
    ctx = pkg.Context(0)
    cfg = sim.config_gauss_small(seed=2, shape=(128, 128), n=40)
    r = pkg.fit_all_stars_simultaneous(cfg["image"], cfg["psf"], cfg["sources"], cfg["fit_rad"], fixed=cfg["fixed"],
                                       inv_var=cfg["inv_var"], model_rad=cfg["model_rad"], max_iter=cfg["max_iter"],
                                       show_trace=True, ctx=ctx)
    lines = [l for l in capsys.readouterr().out.splitlines() if l.startswith("Iter")]
    tr = r.stats["trace"]
    assert len(lines) == len(tr) > 0
    for l, t in zip(lines, tr):
        assert f"Iter {int(t[0]):4d} trial {int(t[1]):2d}" in l and "||g|| =" in l
        assert ("accepted" in l) == (t[3] < t[2])
    assert np.all(tr[:, 7] > 0)   # the gradient cosine of each iteration
