"""Oracle binning, operators, inner solvers, outer loop and lm_irls, pinned
against the reference's own test invariants and closed forms (CPU)."""
import math

import numpy as np
import pytest

import fixtures
import oracle_fit
from _support import f64, golden, load_pkg, oracle, ptr

pkg = load_pkg()
abi = pkg._abi
G = golden()


# --------------------------------------------------------------------- binning
def test_morton_interleave():
    lib = oracle()
    for y, x in [(0, 0), (1, 0), (0, 1), (5, 9), (4095, 17), (70000, 3)]:
        want = 0
        for b in range(32):
            want |= ((y >> b) & 1) << (2 * b)
            want |= ((x >> b) & 1) << (2 * b + 1)
        assert lib.or_morton2d(y, x) == want


def test_ring_offsets_prefix_boxes():
    R = G["ring_offsets"]["R"]
    n = (2 * R + 1) ** 2
    dy = np.zeros(n, np.int32)
    dx = np.zeros(n, np.int32)
    assert oracle().or_ring_offsets(R, ptr(dy, abi.i32p), ptr(dx, abi.i32p)) == n
    for k in range(R + 1):
        m = (2 * k + 1) ** 2
        got = set(zip(dy[:m].tolist(), dx[:m].tolist()))
        assert got == {(a, b) for a in range(-k, k + 1) for b in range(-k, k + 1)}


def _radii(model, mr, nsig, Rf, Rc, w, fl):
    out = np.zeros(len(fl), np.int32)
    w = f64(w)
    fl = f64(fl)
    rc = oracle().or_model_radii(model.desc(), mr, nsig, Rf, Rc, ptr(w, abi.f64p), len(w),
                                 ptr(fl, abi.f64p), len(fl), ptr(out, abi.i32p))
    assert rc == 0
    return out


def test_model_radii_scalar_and_auto():
    g = G["model_radii_scalar"]
    m = pkg.CircularGaussianPSF(fwhm=g["fwhm"])
    w = np.full(g["w_len"], g["w_value"])
    for mr, fl, want in g["cases"]:
        assert list(_radii(m, mr, 1.0, g["R_fit"], g["R_cap"], w, fl)) == want
    a = G["model_radii_auto"]
    rr = _radii(m, 0.0, 1.0, a["R_fit"], a["R_cap"], w, a["fluxes"])
    assert np.all(np.diff(rr) >= 0)
    assert np.all((rr >= a["R_fit"]) & (rr <= a["R_cap"]))
    assert rr[0] == a["R_fit"] and rr[-1] > rr[0]
    assert np.all(_radii(m, 0.0, 3.0, a["R_fit"], a["R_cap"], w, a["fluxes"]) <= rr)


def test_model_radii_auto_closed_form():
    # sigma_bg = 1/sqrt(w); sb(r) from the analytic Gaussian CoG annuli
    m = pkg.CircularGaussianPSF(fwhm=2.5)
    w = np.full(4000, 1 / 125.0)
    Rf, Rc = 2, 20
    F = 5e4
    got = _radii(m, 0.0, 1.0, Rf, Rc, w, [F])[0]
    ee = [1 - math.exp(-4 * math.log(2) * r * r / 6.25) for r in range(1, Rc + 1)]
    ee = [e / ee[-1] for e in ee]
    sb = [max((ee[r - 1] - (ee[r - 2] if r > 1 else 0.0)) / (math.pi * (2 * r - 1)), 0.0) for r in range(1, Rc + 1)]
    r = Rf
    while r < Rc and F * sb[r - 1] > math.sqrt(125.0):
        r += 1
    assert got == r


def build(w, ny, nx, ys, xs, fs, fit_rad):
    lib = oracle()
    n = len(ys)
    R = math.ceil(fit_rad)
    S2 = (2 * R + 1) ** 2
    ay = np.zeros(n, np.int32)
    ax = np.zeros(n, np.int32)
    act = np.zeros(n, np.int32)
    na = np.zeros(1, np.int32)
    pix = np.zeros(S2 * n, np.int32)
    um = np.zeros(ny * nx, np.uint8)
    wv = f64(w.ravel(order="F")) if w is not None else None
    lib.or_build_stamps(ptr(wv, abi.f64p), ny, nx, ptr(f64(ys), abi.f64p), ptr(f64(xs), abi.f64p),
                        ptr(f64(fs), abi.f64p), n, fit_rad, ptr(ay, abi.i32p), ptr(ax, abi.i32p),
                        ptr(act, abi.i32p), ptr(na, abi.i32p), ptr(pix, abi.i32p), ptr(um, abi.u8p))
    k = int(na[0])
    return ay, ax, act[:k], pix[: S2 * k].reshape(k, S2), um


def test_build_stamps_definitions():
    rng = np.random.default_rng(5)
    ny, nx = 40, 50
    n = 60
    ys = rng.uniform(-3, ny + 3, n)
    xs = rng.uniform(-3, nx + 3, n)
    ys[0] = 10.5   # tie -> even (10)
    xs[0] = 11.5   # tie -> even (12)
    ys[1] = np.nan
    fs = rng.uniform(10, 100, n)
    fs[2] = np.inf
    w = np.ones((ny, nx))
    w[5:9, 5:9] = 0
    ay, ax, act, pix, um = build(w, ny, nx, ys, xs, fs, 2.3)
    assert ay[0] == 10 and ax[0] == 12
    R = 3
    # brute force restatement of the definitions
    keys = []
    for i in range(n):
        if not (np.isfinite(ys[i]) and np.isfinite(xs[i])):
            continue
        a, b = int(np.rint(ys[i])), int(np.rint(xs[i]))
        assert (a, b) == (ay[i], ax[i])
        p = []
        for dx in range(-R, R + 1):
            for dy in range(-R, R + 1):
                gy, gx = a + dy, b + dx
                ok = 1 <= gy <= ny and 1 <= gx <= nx and w[gy - 1, gx - 1] > 0
                p.append(gy + (gx - 1) * ny if ok else 0)
        if any(p) and np.isfinite(fs[i]) and a >= 0 and b >= 0:
            keys.append((oracle().or_morton2d(a, b), i, p))
    keys.sort(key=lambda t: (t[0], t[1]))
    assert list(act) == [t[1] for t in keys]
    assert np.array_equal(pix, np.array([t[2] for t in keys], np.int32))
    want = np.zeros(ny * nx, np.uint8)
    for t in keys:
        for fi in t[2]:
            if fi:
                want[fi - 1] = 1
    assert np.array_equal(um, want)


def isapprox(x, y, rtol, atol=0.0):
    """Julia's isapprox on arrays: norm(x - y) <= max(atol, rtol * max(norm(x), norm(y)))."""
    x, y = np.asarray(x, float), np.asarray(y, float)
    return np.linalg.norm(x - y) <= max(atol, rtol * max(np.linalg.norm(x), np.linalg.norm(y)))


# -------------------------------------------------------------------- operators
def _pixel_map(rule_step, p, S2, na, npix):
    pix = np.zeros((na, S2), np.int32)
    for a in range(1, na + 1):
        for m in range(1, S2 + 1):
            pix[a - 1, m - 1] = (m + (a - 1) * rule_step - 1) % npix + 1
    return pix


def _dense(V, pix, p, S2, na, npix):
    J = np.zeros((npix, na * p))
    for a in range(na):
        for m in range(S2):
            fi = pix[a, m]
            if fi:
                J[fi - 1, a * p:(a + 1) * p] = V[a, m]
    return J


def _ops(V, pix, p, S2, na, npix, v=None, u=None, live=None):
    lib = oracle()
    Vf = f64(V.ravel())
    pf = np.ascontiguousarray(pix.ravel(), np.int32)
    lv = None if live is None else np.ascontiguousarray(live, np.uint8)
    y = z = None
    if v is not None:
        y = np.zeros(npix)
        lib.or_apply_J(ptr(y, abi.f64p), npix, ptr(Vf, abi.f64p), ptr(pf, abi.i32p), p, S2, na,
                       ptr(f64(v), abi.f64p), ptr(lv, abi.u8p))
    if u is not None:
        z = np.zeros(na * p)
        lib.or_apply_JT(ptr(z, abi.f64p), ptr(Vf, abi.f64p), ptr(pf, abi.i32p), p, S2, na,
                        ptr(f64(u), abi.f64p), ptr(lv, abi.u8p))
    return y, z


def test_adjoint_identity_and_dense_J():
    c = G["operator_fixtures"]["adjoint"]
    p, S2, na, npix = c["p"], c["S2"], c["n_active"], c["npix"]
    rng = np.random.default_rng(1234)
    V = rng.standard_normal((na, S2, p))
    pix = _pixel_map(2, p, S2, na, npix)
    u = rng.standard_normal(npix)
    v = rng.standard_normal(na * p)
    y, z = _ops(V, pix, p, S2, na, npix, v=v, u=u)
    assert np.dot(u, y) == pytest.approx(np.dot(z, v), rel=c["rtol"])
    J = _dense(V, pix, p, S2, na, npix)
    np.testing.assert_allclose(y, J @ v, rtol=c["rtol"])
    np.testing.assert_allclose(z, J.T @ u, rtol=c["rtol"])
    live = np.ones(na, np.uint8)
    live[1] = 0
    Jm = J.copy()
    Jm[:, p:2 * p] = 0
    y, z = _ops(V, pix, p, S2, na, npix, v=v, u=u, live=live)
    np.testing.assert_allclose(y, Jm @ v, rtol=c["rtol"])
    np.testing.assert_allclose(z, Jm.T @ u, rtol=c["rtol"])
    assert np.all(z[p:2 * p] == 0)


def _equilibrate(V):
    s = np.sqrt((V ** 2).sum(axis=1, keepdims=True))
    return V / np.maximum(s, np.finfo(float).eps)


def _solve(solver, V, pix, p, S2, na, npix, b, lam, itmax, tol, live=None):
    x = np.zeros(na * p)
    lv = np.ones(na, np.uint8) if live is None else np.ascontiguousarray(live, np.uint8)
    its = oracle().or_inner_solve(solver, ptr(f64(V.ravel()), abi.f64p),
                                  ptr(np.ascontiguousarray(pix.ravel(), np.int32), abi.i32p), p, S2, na,
                                  npix, ptr(lv, abi.u8p), ptr(f64(b), abi.f64p), lam, itmax, tol, tol,
                                  ptr(x, abi.f64p))
    return x, its


@pytest.mark.parametrize("solver", [abi.CP_SOLVER_LSQR, abi.CP_SOLVER_CGLS, abi.CP_SOLVER_LSMR])
def test_inner_solve_matches_dense_damped_normal_equations(solver):
    c = G["operator_fixtures"]["lsqr_dense"]
    p, S2, na, npix = c["p"], c["S2"], c["n_active"], c["npix"]
    rng = np.random.default_rng(99)
    V = _equilibrate(rng.standard_normal((na, S2, p)))
    pix = _pixel_map(8, p, S2, na, npix)
    J = _dense(V, pix, p, S2, na, npix)
    mu = c["mu"]
    b = rng.standard_normal(npix)
    want = np.linalg.solve(J.T @ J + mu * np.eye(na * p), J.T @ b)
    x, _ = _solve(solver, V, pix, p, S2, na, npix, b, math.sqrt(mu), c["itmax"], c["tol"])
    assert isapprox(x, want, rtol=c["rtol"], atol=c["atol"])
    live = np.ones(na, np.uint8)
    live[1] = 0
    x, _ = _solve(solver, V, pix, p, S2, na, npix, rng.standard_normal(npix), math.sqrt(mu), c["itmax"],
                  c["tol"], live=live)
    assert np.all(x[p:2 * p] == 0)


@pytest.mark.parametrize("solver", [abi.CP_SOLVER_LSQR, abi.CP_SOLVER_CGLS])
def test_truncated_iterate_is_krylov_subspace_minimiser(solver):
    """x_k minimises ||J x - b||^2 + lam^2 ||x||^2 over K_k(J'J + lam^2 I, J'b):
    pins every truncated inner iteration (itmax = inner_iterations) exactly."""
    c = G["operator_fixtures"]["lsqr_dense"]
    p, S2, na, npix = c["p"], c["S2"], c["n_active"], c["npix"]
    rng = np.random.default_rng(7)
    V = _equilibrate(rng.standard_normal((na, S2, p)))
    pix = _pixel_map(2, p, S2, na, npix)
    J = _dense(V, pix, p, S2, na, npix)
    b = rng.standard_normal(npix)
    lam = 0.1
    H = J.T @ J + lam * lam * np.eye(na * p)
    g = J.T @ b
    for k in range(1, 7):
        K = np.zeros((na * p, k))
        v = g.copy()
        for j in range(k):
            K[:, j] = v
            v = H @ v
        Q, _ = np.linalg.qr(K)
        want = Q @ np.linalg.solve(Q.T @ H @ Q, Q.T @ g)
        x, its = _solve(solver, V, pix, p, S2, na, npix, b, lam, k, 0.0)
        assert its == k
        np.testing.assert_allclose(x, want, rtol=1e-8, atol=1e-10)


def test_lsmr_iterate_minimises_normal_equation_residual():
    """LSMR's x_k minimises ||J'(b - J x) - lam^2 x|| over K_k(J'J + lam^2 I, J'b)
    (Fong & Saunders 2011, s2): pins every truncated LSMR iteration exactly,
    and it differs from the LSQR iterate (which minimises the residual)."""
    c = G["operator_fixtures"]["lsqr_dense"]
    p, S2, na, npix = c["p"], c["S2"], c["n_active"], c["npix"]
    rng = np.random.default_rng(11)
    V = _equilibrate(rng.standard_normal((na, S2, p)))
    pix = _pixel_map(2, p, S2, na, npix)
    J = _dense(V, pix, p, S2, na, npix)
    b = rng.standard_normal(npix)
    lam = 0.1
    H = J.T @ J + lam * lam * np.eye(na * p)
    g = J.T @ b
    for k in range(1, 7):
        K = np.zeros((na * p, k))
        v = g.copy()
        for j in range(k):
            K[:, j] = v
            v = H @ v
        Q, _ = np.linalg.qr(K)
        y = np.linalg.lstsq(H @ Q, g, rcond=None)[0]
        want = Q @ y
        x, its = _solve(abi.CP_SOLVER_LSMR, V, pix, p, S2, na, npix, b, lam, k, 0.0)
        assert its == k
        np.testing.assert_allclose(x, want, rtol=1e-8, atol=1e-10)
        xq, _ = _solve(abi.CP_SOLVER_LSQR, V, pix, p, S2, na, npix, b, lam, k, 0.0)
        if k > 1:
            assert np.max(np.abs(xq - x)) > 1e-6


@pytest.mark.parametrize("solver", ["lsqr", "lsmr"])
def test_inner_solvers_match_scipy_iterate_by_iterate(solver):
    """The oracle's LSQR / LSMR against scipy.sparse.linalg's implementations of
    the same published methods (damp = lam, x0 = 0) at every truncated
    iteration count, on the stamp operator written out as a dense matrix."""
    sla = pytest.importorskip("scipy.sparse.linalg")
    c = G["operator_fixtures"]["lsqr_dense"]
    p, S2, na, npix = c["p"], c["S2"], c["n_active"], c["npix"]
    rng = np.random.default_rng(21)
    V = _equilibrate(rng.standard_normal((na, S2, p)))
    pix = _pixel_map(3, p, S2, na, npix)
    J = _dense(V, pix, p, S2, na, npix)
    b = rng.standard_normal(npix)
    lam = 0.05
    code = abi.CP_SOLVER_LSQR if solver == "lsqr" else abi.CP_SOLVER_LSMR
    fn = sla.lsqr if solver == "lsqr" else sla.lsmr
    for k in range(1, 5):   # below the 5-deep forward-error window scipy does not have
        want = fn(J, b, damp=lam, atol=0.0, btol=0.0, conlim=0.0, iter_lim=k)[0] if solver == "lsqr" else \
            fn(J, b, damp=lam, atol=0.0, btol=0.0, conlim=0.0, maxiter=k)[0]
        x, its = _solve(code, V, pix, p, S2, na, npix, b, lam, k, 0.0)
        assert its == k
        np.testing.assert_allclose(x, want, rtol=1e-10, atol=1e-12)


def test_cg_iterates_equal_lsqr_in_exact_arithmetic_and_stop_rule():
    """CG on the damped normal equations and LSQR span the same Krylov space and
    minimise the same quadratic, so their k-th iterates agree to rounding; CG
    stops at the first k with ||s_k|| <= atol ||s_0||, which the dense normal
    equations give independently."""
    c = G["operator_fixtures"]["lsqr_dense"]
    p, S2, na, npix = c["p"], c["S2"], c["n_active"], c["npix"]
    rng = np.random.default_rng(31)
    V = _equilibrate(rng.standard_normal((na, S2, p)))
    pix = _pixel_map(2, p, S2, na, npix)
    J = _dense(V, pix, p, S2, na, npix)
    b = rng.standard_normal(npix)
    lam = 0.2
    for k in range(1, 8):
        xc, _ = _solve(abi.CP_SOLVER_CGLS, V, pix, p, S2, na, npix, b, lam, k, 0.0)
        xq, _ = _solve(abi.CP_SOLVER_LSQR, V, pix, p, S2, na, npix, b, lam, k, 0.0)
        np.testing.assert_allclose(xc, xq, rtol=1e-8, atol=1e-10)
    H = J.T @ J + lam * lam * np.eye(na * p)
    g = J.T @ b
    atol = 1e-3
    _, its = _solve(abi.CP_SOLVER_CGLS, V, pix, p, S2, na, npix, b, lam, 200, atol)
    for k in range(1, its + 1):
        xk, _ = _solve(abi.CP_SOLVER_CGLS, V, pix, p, S2, na, npix, b, lam, k, 0.0)
        r = np.linalg.norm(g - H @ xk)
        assert (r <= atol * np.linalg.norm(g) * (1 + 1e-9)) == (k == its), (k, its, r)


# ------------------------------------------------------------------- whole fit
def test_isolated_noiseless_recovery():
    g = G["isolated_recovery"]
    rng = np.random.default_rng(42)
    shape = tuple(g["shape"])
    ys, xs = fixtures.separated_positions(rng, g["n"], shape, 10, 15)
    fs = rng.uniform(*g["flux_range"], g["n"])
    img = fixtures.render_gauss(shape, ys, xs, fs, g["fwhm"], 5)
    psf = pkg.CircularGaussianPSF(fwhm=g["fwhm"])
    for solver in (abi.CP_SOLVER_LSQR, abi.CP_SOLVER_CGLS):
        r = oracle_fit.fit(psf, img, ys, xs, np.full(g["n"], g["init_flux"]), g["fit_rad"],
                           model_rad=float(g["model_rad"]), max_iter=g["max_iter"], x_tol=g["x_tol"],
                           solver=solver)
        assert r.valid.all()
        np.testing.assert_allclose(r.flux, fs, rtol=g["flux_rtol"])
        np.testing.assert_allclose(r.y, ys, atol=1e-10)
        np.testing.assert_allclose(r.x, xs, atol=1e-10)


def test_isolated_noiseless_recovery_epsf():
    """Same fixture through a one-node GriddedPSFModel{ImagePSF} (the reference
    loops the isolated test over both psf types, :223-253)."""
    rng = np.random.default_rng(42)
    shape = (128, 128)
    ys, xs = fixtures.separated_positions(rng, 5, shape, 10, 15)
    fs = rng.uniform(600, 900, 5)
    epsf = pkg.moffat_epsf(fwhm=2.2, beta=2.5, size=41, oversampling=4)
    gpsf = pkg.GriddedPSFModel([epsf], [0.0], [0.0])
    img = np.zeros(shape)
    lib = oracle()
    d = gpsf.desc()
    for y, x, f in zip(ys, xs, fs):
        for gy in range(math.floor(y - 5), math.ceil(y + 5) + 1):
            for gx in range(math.floor(x - 5), math.ceil(x + 5) + 1):
                img[gy - 1, gx - 1] += lib.or_evaluate(d, ptr(f64([y, x, f, 0.0]), abi.f64p), gy, gx)
    r = oracle_fit.fit(gpsf, img, ys, xs, np.full(5, 400.0), 5, model_rad=5.0, max_iter=40, x_tol=1e-8)
    assert r.valid.all()
    np.testing.assert_allclose(r.flux, fs, rtol=1e-10)
    np.testing.assert_allclose(r.y, ys, atol=1e-10)


def _crowded():
    img, ys, xs, fs = fixtures.crowded_field(11, min_sep=2.0)
    return img, ys, xs, fs, pkg.CircularGaussianPSF(fwhm=2.0)


@pytest.mark.parametrize("other", [abi.CP_SOLVER_LSMR, abi.CP_SOLVER_CGLS])
def test_crowded_lsqr_agrees_with_other_solver_and_freezes(other):
    """The LSQR-vs-LSMR crowded fixture (:255-292), and the same bars for CG."""
    img, ys, xs, fs, psf = _crowded()
    kw = dict(max_iter=25, inner_iterations=10, x_tol=1e-8, f_tol=0.0, model_rad=5.0)
    a = oracle_fit.fit(psf, img, ys, xs, fs.copy(), 5, **kw)
    b = oracle_fit.fit(psf, img, ys, xs, fs.copy(), 5, solver=other, **kw)
    assert np.sum(a.n_iter[a.valid] < a.n_passes) > 0.5 * a.valid.sum()
    g = a.valid & b.valid
    assert g.sum() > 0.8 * len(g)
    # Julia's isapprox on vectors: a norm-wise relative bar
    assert np.linalg.norm(a.flux[g] - b.flux[g]) <= 1e-2 * max(np.linalg.norm(a.flux[g]), np.linalg.norm(b.flux[g]))
    for d in (np.abs(a.y[g] - b.y[g]), np.abs(a.x[g] - b.x[g])):
        assert np.median(d) < 1e-5 and np.quantile(d, 0.9) < 1e-3 and d.max() < 2e-2
    assert not np.array_equal(a.flux, b.flux)   # the solvers are different


def test_frozen_stars_stable_under_extra_iterations():
    """:426-453 -- stars frozen by iteration K1 take an identical path."""
    rng = np.random.default_rng(99)
    shape = (200, 200)
    n = 60
    ys = rng.uniform(10, 190, n)
    xs = rng.uniform(10, 190, n)
    fs = (200.0 ** -1 + rng.random(n) * (3000.0 ** -1 - 200.0 ** -1)) ** -1
    model = fixtures.render_gauss(shape, ys, xs, fs, 2.0, 6) + 50
    img = rng.poisson(model) + rng.normal(0, 2.0, shape) - 50
    psf = pkg.CircularGaussianPSF(fwhm=2.0)
    r1 = oracle_fit.fit(psf, img, ys, xs, fs, 5, max_iter=10, model_rad=5.0)
    r2 = oracle_fit.fit(psf, img, ys, xs, fs, 5, max_iter=25, model_rad=5.0)
    early = [i for i in range(n) if r1.valid[i] and r2.valid[i] and 0 < r2.n_iter[i] <= 10]
    assert len(early) > 0
    for i in early:
        assert r1.y[i] == r2.y[i] and r1.x[i] == r2.x[i] and r1.flux[i] == r2.flux[i]
    g = r2.valid
    assert np.any(r2.n_iter[g] < r2.n_passes) and len(np.unique(r2.n_iter[g])) > 1


def test_contract_errors_and_edge_cases():
    psf = pkg.CircularGaussianPSF(fwhm=2.0)
    img = np.zeros((32, 32))
    # empty catalogue
    r = oracle_fit.fit(psf, np.zeros((16, 16)), [], [], [], 5, model_rad=5.0)
    assert r.n_passes == 0
    # non-finite image without inv_var
    bad = img.copy()
    bad[4, 4] = np.nan
    with pytest.raises(ValueError, match="non-finite"):
        oracle_fit.fit(psf, bad, [16.0], [16.0], [100.0], 5, model_rad=5.0)
    # fully masked star -> excluded
    iv = np.ones((32, 32))
    iv[9:22, 9:22] = 0
    r = oracle_fit.fit(psf, img, [16.0], [16.0], [100.0], 5, inv_var=iv)
    assert r.n_failed == 1 and not r.valid[0]
    # :auto needs inv_var
    with pytest.raises(ValueError, match="requires inv_var"):
        oracle_fit.fit(psf, np.zeros((40, 40)), [20.0], [20.0], [100.0], 3)
    r = oracle_fit.fit(psf, np.zeros((40, 40)), [20.0], [20.0], [100.0], 3, model_rad=3.0, max_iter=5)
    assert len(r.flux) == 1
    for kw in [dict(max_iter=0), dict(inner_iterations=0), dict(linear_tol=0.0), dict(max_trials=0),
               dict(model_rad_nsigma=0.0), dict(model_rad_max=1.0)]:
        with pytest.raises(ValueError):
            oracle_fit.fit(psf, img, [16.0], [16.0], [100.0], 5, model_rad=5.0, **kw)


def test_wing_truncation_bias_removed_by_model_rad():
    """:499-528 -- a single fit radius biases a bright star ringed by faint
    neighbours; a wide or :auto model_rad removes the bias.  The reference's
    exact neighbour draw comes from StableRNG (unavailable), so the property
    is checked over several seeded layouts of the same construction."""
    psf = pkg.CircularGaussianPSF(fwhm=2.5)
    area = math.pi * 2.5 ** 2 / (2 * math.log(2))  # effective_area, parametric_models.jl:83

    def ffs(s):  # flux_for_snr(background=100, read_noise=5, gain=1), simulation.jl:59-80
        return (s * s + math.sqrt(s ** 4 + 4 * s * s * area * 125.0)) / 2

    singles = []
    for seed in range(4):
        rng = np.random.default_rng(seed)
        ys = [25.35] + list(6 + 38 * rng.random(40))
        xs = [25.65] + list(6 + 38 * rng.random(40))
        fs = [ffs(300.0)] + [ffs(3 + 9 * v) for v in rng.random(40)]
        img = fixtures.render_gauss((50, 50), ys, xs, fs, 2.5, 6.25)
        iv = 1.0 / (100.0 + img + 25.0)
        kw = dict(inv_var=iv, inner_iterations=30, max_iter=80)
        bias = [(oracle_fit.fit(psf, img, ys, xs, fs, 2, model_rad=mr, **kw).flux[0] - fs[0]) / fs[0]
                for mr in (2.0, 0.0, 8.0)]
        assert abs(bias[1]) < 1e-3 and abs(bias[2]) < 1e-3
        assert abs(bias[0]) > 10 * max(abs(bias[1]), abs(bias[2]))
        singles.append(abs(bias[0]))
    assert max(singles) > 5e-3


# ------------------------------------------------------------------- lm_irls
def test_lm_line_fit():
    g = G["lm_line"]
    xs = f64(g["xs"])
    ys = f64([g["slope"] * v + g["intercept"] for v in g["xs"]])
    o = abi.lm_opts_default()
    o.max_iter, o.x_tol, o.f_tol, o.g_tol = g["max_iter"], g["x_tol"], g["f_tol"], g["g_tol"]
    x = np.zeros(2)
    st = abi.LMStat()
    import ctypes as C
    assert oracle().or_lm_line(ptr(xs, abi.f64p), ptr(ys, abi.f64p), len(xs), ptr(f64(g["x0"]), abi.f64p),
                               C.byref(o), ptr(x, abi.f64p), C.byref(st)) == 0
    assert st.converged
    np.testing.assert_allclose(x, g["expect_minimizer"], rtol=g["rtol"])
    assert st.minimum < g["max_cost"]


def _fit_star(model, img, theta0, mask, iv=None, **kw):
    import ctypes as C
    o = abi.lm_opts_default()
    for k, v in kw.items():
        setattr(o, k, v)
    ny, nx = img.shape
    npf = abi.NPARAMS[model.kind]
    p = bin(mask).count("1")
    out = np.zeros(npf)
    cov = np.zeros(p * p)
    st = abi.LMStat()
    ivf = None if iv is None else f64(iv.ravel(order="F"))
    rc = oracle().or_fit_star(model.desc(), ptr(f64(img.ravel(order="F")), abi.f64p), ny, nx,
                              ptr(ivf, abi.f64p), ptr(f64(theta0), abi.f64p), mask, C.byref(o),
                              ptr(out, abi.f64p), ptr(cov, abi.f64p), C.byref(st))
    return rc, out, cov.reshape(p, p), st


def test_fit_star_gaussian_noiseless_recovery():
    """test/psf/psf_fitting_tests.jl:59-86."""
    truth = pkg.CircularGaussianPSF(x=13.5, y=12.3, fwhm=4.0, flux=200.0, bkg=5.0)
    ii, jj = np.meshgrid(np.arange(1, 31), np.arange(1, 31), indexing="ij")
    lib = oracle()
    d = truth.desc()
    tp = f64(truth.params())
    img = np.array([[lib.or_evaluate(d, ptr(tp, abi.f64p), float(i), float(j)) for j in range(1, 31)]
                    for i in range(1, 31)])
    init = [11.8, 14.1, 4.3, 190.0, 5.5]
    rc, out, cov, st = _fit_star(truth, img, init, 0b11111, x_tol=1e-6)
    assert rc == 0 and st.converged
    np.testing.assert_allclose(out, truth.params(), rtol=1e-3)
    rc, out, cov, st = _fit_star(truth, img, init, 0b11111, iv=np.ones_like(img), x_tol=1e-6)
    assert cov.shape == (5, 5) and rc == 0
    rc, out, _, _ = _fit_star(truth, img, [11.8, 14.1, 4.3, 190.0, 5.0], 0b01111, x_tol=1e-6)
    assert out[4] == 5.0
    np.testing.assert_allclose(out[1], 13.5, rtol=1e-3)
    # all masked -> too few pixels (ArgumentError)
    rc, *_ = _fit_star(truth, img, init, 0b11111, iv=np.zeros_like(img))
    assert rc == abi.CP_EARG


def test_fit_star_moffat_noiseless_recovery():
    truth = pkg.CircularMoffatPSF(y=8.2, x=7.7, alpha=2.1, beta=2.5, flux=900.0, bkg=40.0)
    lib = oracle()
    d = truth.desc()
    tp = f64(truth.params())
    img = np.array([[lib.or_evaluate(d, ptr(tp, abi.f64p), float(i), float(j)) for j in range(1, 16)]
                    for i in range(1, 16)])
    rc, out, cov, st = _fit_star(truth, img, [8.5, 7.4, 2.1, 2.5, 800.0, 44.0], 0b110011)
    assert rc == 0 and st.converged
    np.testing.assert_allclose(out, truth.params(), rtol=1e-6)
