"""The oracle's outer Gauss-Newton / LM control flow pinned trial by trial
against a dense numpy restatement (tests/dense_lm.py) of
psf_photometry_simultaneous.jl:1186-1284, :700-716 and :739-823 on a few
blended Gaussian stars.

With inner_iterations well above the 3N unknowns and a tiny linear_tol, the
oracle's LSQR reaches the damped normal-equation solve that the restatement
computes densely, so the two must take the same decisions: for every trial
the same (iteration, trial, live count, lambda) and the same cost and
candidate cost (hence the same accept / reject); for every star the same
freeze iteration and convergence flag.  The fixtures are chosen to exercise
rejected trials with lambda_up != lambda_down, per-star x and g freezing and
the f test, and each planted mistake of the restatement changes the
comparison, so a wrong lambda update, f test or freeze rule in the oracle
would fail here."""
import numpy as np
import pytest

import dense_lm
import fixtures
import oracle_fit
from _support import load_pkg

pkg = load_pkg()
PSF = pkg.CircularGaussianPSF(fwhm=2.5)


def _fixture(seed, n=5):
    rng = np.random.default_rng(seed)
    shape = (48, 48)
    ys, xs, fs = rng.uniform(10, 38, n), rng.uniform(10, 38, n), rng.uniform(500, 6000, n)
    img = fixtures.render_gauss(shape, ys, xs, fs, 2.5, 8) + rng.normal(0, 3.0, shape)
    iv = np.full(shape, 1 / 9.0)
    y0 = ys + rng.normal(0, 0.8, n)
    x0 = xs + rng.normal(0, 0.8, n)
    f0 = fs * rng.uniform(0.3, 2.5, n)
    return img, iv, y0, x0, f0


CASES = {
    # rejected trials, per-star x freezing and an f-test finish
    "x_freeze": dict(seed=21, opts=dict(max_iter=60, lambda_down=4.0, lambda_init=1e-6)),
    # x test disabled: per-star g freezing, rejected trials, f-test finish
    "g_freeze": dict(seed=21, opts=dict(max_iter=60, lambda_down=4.0, lambda_init=1e-6, x_tol=1e-12)),
}


def _oracle(img, iv, y0, x0, f0, opts):
    return oracle_fit.fit(PSF, img, y0, x0, f0, 4, inv_var=iv, model_rad=8.0, trace=2000, inner_iterations=60,
                          linear_tol=1e-14, **opts)


def _same(ro, d):
    t, td = ro.trace, d["trace"]
    return (len(t) == len(td) and np.array_equal(t[:, [0, 1, 5]], td[:, [0, 1, 5]])
            and np.allclose(t[:, 2:5], td[:, 2:5], rtol=1e-7, atol=0)
            and np.array_equal(ro.n_iter, d["n_iter"]) and np.array_equal(ro.converged.astype(bool), d["converged"])
            and ro.n_passes == d["n_passes"])


@pytest.mark.parametrize("case", sorted(CASES))
def test_oracle_trial_trace_equals_dense_restatement(case):
    c = CASES[case]
    img, iv, y0, x0, f0 = _fixture(c["seed"])
    ro = _oracle(img, iv, y0, x0, f0, c["opts"])
    d = dense_lm.dense_fit(img, iv, y0, x0, f0, 2.5, 4, 8.0, **c["opts"])
    # the fixture's decisions are not at a rounding-level threshold
    ro2 = _oracle(img * (1 + 1e-12), iv, y0, x0, f0, c["opts"])
    assert np.array_equal(ro.trace[:, [0, 1, 5]], ro2.trace[:, [0, 1, 5]]) and np.array_equal(ro.n_iter, ro2.n_iter)
    t = ro.trace
    assert np.sum(t[:, 3] >= t[:, 2]) >= 3, "the fixture has too few rejected trials"
    assert len(np.unique(ro.n_iter)) >= 3, "the stars should freeze at different iterations"
    assert _same(ro, d), (t[:, :6], d["trace"])
    np.testing.assert_allclose(ro.flux, d["theta"][:, 2], rtol=1e-7)
    np.testing.assert_allclose(ro.y, d["theta"][:, 0], rtol=0, atol=1e-7)
    np.testing.assert_allclose(ro.x, d["theta"][:, 1], rtol=0, atol=1e-7)


@pytest.mark.parametrize("variant", ["lambda_up", "f_test", "x_freeze", "g_freeze"])
def test_each_planted_mistake_is_caught(variant):
    caught = []
    for case in sorted(CASES):
        c = CASES[case]
        img, iv, y0, x0, f0 = _fixture(c["seed"])
        ro = _oracle(img, iv, y0, x0, f0, c["opts"])
        dv = dense_lm.dense_fit(img, iv, y0, x0, f0, 2.5, 4, 8.0, variant=variant, **c["opts"])
        caught.append(not _same(ro, dv))
    assert any(caught), f"a {variant} mistake would go unnoticed"
