"""Python wrappers around the oracle's simultaneous fit (checker only)."""
import ctypes as C

import numpy as np

from _support import f64, load_pkg, oracle, ptr

abi = load_pkg()._abi


FIELDS = ("y", "x", "flux", "bkg", "y_err", "x_err", "flux_err", "bkg_err", "chisq",
          "qfit", "qfit_expected", "qfit_z", "crowding", "spread_model", "spread_model_err")


class Result:
    pass


def opts(**kw):
    o = abi.sim_opts_default()
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def fit(psf_model, image, ys, xs, fs, fit_rad, inv_var=None, trace=0, **kw):
    """image: numpy [ny, nx] (row = y).  Returns a Result with numpy fields."""
    lib = oracle()
    ny, nx = image.shape
    n = len(ys)
    img = f64(image.ravel(order="F"))
    iv = None if inv_var is None else f64(np.asarray(inv_var).ravel(order="F"))
    r = Result()
    for k in FIELDS:
        setattr(r, k, np.zeros(n))
    r.converged = np.zeros(n, np.uint8)
    r.valid = np.zeros(n, np.uint8)
    r.n_iter = np.zeros(n, np.int32)
    r.active = np.zeros(n, np.uint8)
    r.residual = np.zeros(ny * nx, np.float32)
    r.trace = np.zeros(max(trace, 1) * 8)
    s = abi.SimResult()
    for k in FIELDS:
        setattr(s, k, ptr(getattr(r, k), abi.f64p))
    s.converged = ptr(r.converged, abi.u8p)
    s.valid = ptr(r.valid, abi.u8p)
    s.n_iter = ptr(r.n_iter, abi.i32p)
    s.active = ptr(r.active, abi.u8p)
    s.residual = ptr(r.residual, abi.f32p)
    s.trace = ptr(r.trace, abi.f64p)
    s.trace_cap = trace
    o = opts(**kw)
    d = psf_model.desc()
    rc = lib.or_fit_simultaneous(d, ptr(img, abi.f64p), ptr(iv, abi.f64p), ny, nx, ptr(f64(ys), abi.f64p),
                                 ptr(f64(xs), abi.f64p), ptr(f64(fs), abi.f64p), n, float(fit_rad),
                                 C.byref(o), C.byref(s))
    if rc != 0:
        raise ValueError(lib.or_last_error().decode())
    r.n_passes, r.n_failed, r.n_active, r.n_union = s.n_passes, s.n_failed, s.n_active, s.n_union
    r.cost_final = s.cost_final
    r.trace = r.trace[: 8 * s.trace_len].reshape(-1, 8)
    r.residual = r.residual.reshape(nx, ny).T
    r.valid = r.valid.astype(bool)
    r.converged = r.converged.astype(bool)
    return r


def fit_all(psf_model, image, sources, fit_rad, n_passes=3, fixed=None, inv_var=None, spread_model_fwhm=None,
            **lm_kw):
    """Oracle fit_all_stars with the package function's arguments: image
    [ny, nx] (row = y), sources a dict with y, x and optional flux, bkg."""
    lib = oracle()
    fixed = dict(fixed or {})
    ny, nx = np.shape(image)
    names = psf_model.props
    y = f64(sources["y"])
    n = len(y)
    cols = {"y": y, "x": f64(sources["x"]), "flux": f64(sources.get("flux", np.ones(n))),
            "bkg": f64(sources.get("bkg", np.zeros(n)))}
    th0 = np.tile(psf_model.params(), (max(n, 1), 1))[:n].copy()
    for k, nme in enumerate(names):
        if nme in cols:
            th0[:, k] = cols[nme]
    th0 = f64(th0)
    mask = sum(1 << k for k, nme in enumerate(names) if nme not in fixed)
    fixv = f64([float(fixed[k]) if k in fixed else np.nan for k in names])
    o = abi.lm_opts_default()
    for k, v in lm_kw.items():
        setattr(o, k, v)
    r = Result()
    for k in FIELDS:
        setattr(r, k, np.zeros(n))
    r.converged = np.zeros(n, np.uint8)
    r.valid = np.zeros(n, np.uint8)
    r.n_iter = np.zeros(n, np.int32)
    r.residual = np.zeros(ny * nx, np.float32)
    s = abi.SimResult()
    for k in FIELDS:
        setattr(s, k, ptr(getattr(r, k), abi.f64p))
    s.converged = ptr(r.converged, abi.u8p)
    s.valid = ptr(r.valid, abi.u8p)
    s.n_iter = ptr(r.n_iter, abi.i32p)
    s.residual = ptr(r.residual, abi.f32p)
    img = f64(np.asarray(image).ravel(order="F"))
    iv = None if inv_var is None else f64(np.asarray(inv_var).ravel(order="F"))
    rc = lib.or_fit_all_stars(psf_model.desc(), ptr(img, abi.f64p), ptr(iv, abi.f64p), ny, nx, ptr(th0, abi.f64p),
                              n, float(fit_rad), int(n_passes), mask, ptr(fixv, abi.f64p), C.byref(o),
                              0.0 if spread_model_fwhm is None else float(spread_model_fwhm), C.byref(s))
    if rc != 0:
        raise ValueError(lib.or_last_error().decode())
    r.n_passes, r.n_failed = s.n_passes, s.n_failed
    r.residual = r.residual.reshape(nx, ny).T
    r.valid = r.valid.astype(bool)
    r.converged = r.converged.astype(bool)
    return r
