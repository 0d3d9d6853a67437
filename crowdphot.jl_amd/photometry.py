"""Host mirror of CrowdPhot.jl's fitting entry points over libcrowdphot_hip.so.

  fit_all_stars_simultaneous   src/photometry/psf_photometry_simultaneous.jl:953-1431
  fit_all_stars                src/photometry/psf_photometry_single.jl:267-469
  MultiPassPhotResult          src/photometry/psf_photometry_single.jl:99-122
  fit_star / LMResult          src/psf/psf_fitting.jl:229-308, src/levenberg_marquardt.jl:28-41
  fit_star_batch               many fit_star calls as one batched GPU launch

Images are numpy arrays indexed [row = y, column = x] with 1-based pixel
centres, as in the reference; they are handed to the library column-major.
Argument errors raise ArgumentError (a ValueError), like the reference's
ArgumentError; device failures raise RuntimeError.  Every computation runs on
the GPU: there is no CPU path.
"""
import ctypes as C
import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _abi
from .psf import CircularGaussianPSF, CircularMoffatPSF, GriddedPSFModel, ImagePSF


class ArgumentError(ValueError):
    """Contract violation (the reference's ArgumentError)."""


class Context:
    """One HIP device context (stream + HBM budget) of libcrowdphot_hip.so."""

    def __init__(self, device=0, hbm_budget=0):
        self.lib = _abi.load_library()
        h = C.c_void_p()
        rc = self.lib.cp_create(C.byref(h), int(device), int(hbm_budget))
        if rc != 0:
            raise RuntimeError(f"cp_create failed ({rc}): no usable HIP device {device}")
        self.handle = h

    def check(self, rc):
        if rc == 0:
            return
        msg = self.lib.cp_last_error(self.handle).decode()
        if rc == _abi.CP_EARG:
            raise ArgumentError(msg)
        if rc == _abi.CP_ENOMEM:
            raise MemoryError(msg)
        raise RuntimeError(f"libcrowdphot_hip error {rc}: {msg}")

    @property
    def stream(self):
        return self.lib.cp_stream(self.handle)

    def get_tuning(self):
        """The context's launch shapes (cp_tuning) as a dict; 0 = built-in default."""
        t = _abi.Tuning()
        self.check(self.lib.cp_get_tuning(self.handle, C.byref(t)))
        return {name: int(getattr(t, name)) for name, _ in _abi.Tuning._fields_}

    def set_tuning(self, **shape):
        """Set launch shapes by cp_tuning field name (blocks per CU, 0 = default);
        fields not given keep their current value."""
        cur = self.get_tuning()
        unknown = set(shape) - set(cur)
        if unknown:
            raise ArgumentError(f"unknown tuning fields {sorted(unknown)}; known: {sorted(cur)}")
        cur.update({k: int(v) for k, v in shape.items()})
        self.check(self.lib.cp_set_tuning(self.handle, C.byref(_abi.Tuning(**cur))))

    def mem_stats(self):
        """Device-block accounting of the context (cp_mem_stats): budget, in_use,
        cached and peak bytes."""
        m = _abi.MemStats()
        self.check(self.lib.cp_mem_stats_get(self.handle, C.byref(m)))
        return {name: int(getattr(m, name)) for name, _ in _abi.MemStats._fields_}

    def set_budget(self, hbm_budget):
        """Cap the context's device blocks at hbm_budget bytes (0 = no cap); restarts the peak."""
        self.check(self.lib.cp_set_budget(self.handle, int(hbm_budget)))

    def close(self):
        if getattr(self, "handle", None):
            self.lib.cp_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


def _rowmajor32(img):
    """float32 C-order view or copy: the library transposes it on the device
    (CP_LAYOUT_ROWMAJOR), so no host-side transpose."""
    a = np.ascontiguousarray(img, dtype=np.float32)
    if a.ndim != 2:
        raise ArgumentError("image must be a 2-D array")
    return a.ravel()


def _colmajor32(img):
    a = np.asarray(img, dtype=np.float32)
    if a.ndim != 2:
        raise ArgumentError("image must be a 2-D array")
    return np.ascontiguousarray(a.T).ravel()  # flat index (gy-1) + (gx-1)*ny


@dataclass
class MultiPassPhotResult:
    y: np.ndarray
    x: np.ndarray
    y_err: np.ndarray
    x_err: np.ndarray
    flux: np.ndarray
    flux_err: np.ndarray
    bkg: np.ndarray
    bkg_err: np.ndarray
    converged: np.ndarray
    valid: np.ndarray
    chisq: np.ndarray
    qfit: np.ndarray
    qfit_expected: np.ndarray
    qfit_z: np.ndarray
    crowding: np.ndarray
    spread_model: np.ndarray
    spread_model_err: np.ndarray
    n_iter: np.ndarray
    n_passes: int
    n_failed: int
    failure_msgs: List[str]
    residual: Optional[np.ndarray]
    stats: dict = field(default_factory=dict)


def _field(rec, name):
    return rec[name] if isinstance(rec, dict) else getattr(rec, name)


def _catalog(sources):
    """_extract_source_catalog (psf_photometry_single.jl:152-197): a
    MatchedFilterResult gives the peak pixels and peak fluxes; a sequence of
    per-star records (the Vector{NamedTuple} of measure_star_shapes) gives
    centroid.y, centroid.x and matched_filter_flux; a NamedTuple-like table
    gives y, x (required) and flux (default 1)."""
    if hasattr(sources, "peaks") and hasattr(sources, "peak_fluxes"):
        pk = np.asarray(sources.peaks, dtype=np.float64).reshape(-1, 2)
        return (np.ascontiguousarray(pk[:, 0]), np.ascontiguousarray(pk[:, 1]),
                np.ascontiguousarray(sources.peak_fluxes, dtype=np.float64))
    if isinstance(sources, (list, tuple)) and not (isinstance(sources, tuple) and hasattr(sources, "_fields")):
        try:
            y = [float(_field(_field(s, "centroid"), "y")) for s in sources]
            x = [float(_field(_field(s, "centroid"), "x")) for s in sources]
            f = [float(_field(s, "matched_filter_flux")) for s in sources]
        except (KeyError, AttributeError, TypeError) as e:
            raise ArgumentError("a per-star source list needs centroid.y, centroid.x and matched_filter_flux "
                                f"in every record ({e})") from None
        return (np.asarray(y, np.float64), np.asarray(x, np.float64), np.asarray(f, np.float64))
    get = (lambda k: sources[k]) if isinstance(sources, dict) else (lambda k: getattr(sources, k))
    has = (lambda k: k in sources) if isinstance(sources, dict) else (lambda k: hasattr(sources, k))
    y = np.ascontiguousarray(get("y"), dtype=np.float64)
    x = np.ascontiguousarray(get("x"), dtype=np.float64)
    if y.shape != x.shape or y.ndim != 1:
        raise ArgumentError("sources.y and sources.x must be 1-D of equal length")
    flux = (np.ascontiguousarray(get("flux"), dtype=np.float64) if has("flux")
            else np.ones_like(y))
    return y, x, flux


def _sim_psf_and_mask(psf, fixed):
    """Free set in getproperties order (psf_fitting.jl:8-14) and the contract
    that only (y, x, flux) may be free (psf_photometry_simultaneous.jl:1022-1031)."""
    names = psf.props
    free = [k for k in names if k not in fixed]
    if not free:
        raise ArgumentError("all model parameters are fixed; nothing to fit")
    offenders = [k for k in free if k not in ("y", "x", "flux")]
    if offenders:
        raise ArgumentError(
            f"fit_all_stars_simultaneous fits only (y, x, flux) per star; got free parameters {offenders}. "
            "Fix `bkg` (subtract the background first and pass `fixed = (; bkg = zero(eltype(image)))`), "
            "and fix all shape parameters.")
    shape_fixed = {k: v for k, v in fixed.items() if k in names and k not in ("y", "x", "flux", "bkg")}
    model = psf.setproperties(**shape_fixed) if shape_fixed else psf
    mask = (_abi.CP_FREE_Y if "y" in free else 0) | (_abi.CP_FREE_X if "x" in free else 0) | \
        (_abi.CP_FREE_FLUX if "flux" in free else 0)
    return model, mask


class _SimCall:
    """Validated inputs, options and caller-owned outputs of one simultaneous fit
    (argument contract of psf_photometry_simultaneous.jl:984-1040,1060-1066)."""

    def __init__(self, image, psf, sources, fit_rad, *, fixed=None, inv_var=None, spread_model_fwhm=None,
                 solver="lsqr", inner_iterations=10, linear_tol=1e-6, model_rad="auto", model_rad_max=None,
                 model_rad_nsigma=1.0, max_step=0.25, max_trials=8, max_iter=40, x_tol=1e-5, f_tol=1e-4,
                 g_tol=1e-4, lambda_init=1e-4, lambda_up=10.0, lambda_down=10.0, lambda_min=1e-12,
                 lambda_max=1e12, damping="marquardt", show_trace=False, covariance_estimator=None,
                 precision="mixed", residual=True, trace=0, tile=0):
        fixed = dict(fixed or {})
        if max_iter <= 0:
            raise ArgumentError("max_iter must be positive")
        if inner_iterations <= 0:
            raise ArgumentError("inner_iterations must be positive")
        if not linear_tol > 0:
            raise ArgumentError("linear_tol must be positive")
        if max_trials <= 0:
            raise ArgumentError("max_trials must be positive")
        if not (model_rad == "auto" or (np.isscalar(model_rad) and model_rad > 0)):
            raise ArgumentError(f"model_rad must be :auto or a positive real, got {model_rad!r}")
        if model_rad_max is None:
            model_rad_max = 10 * fit_rad
        if not model_rad_max >= fit_rad:
            raise ArgumentError("model_rad_max must be >= fit_rad")
        if not model_rad_nsigma > 0:
            raise ArgumentError("model_rad_nsigma must be positive")
        if solver not in ("lsqr", "lsmr", "cg"):
            raise ArgumentError(f"solver must be :lsqr or :lsmr (or the build's :cg), got {solver!r}")
        if damping == "levenberg":
            raise ArgumentError("LevenbergDamping is not supported by fit_all_stars_simultaneous: it conflicts "
                                "with the column-equilibrated x_tol scaling. Use MarquardtDamping.")
        if not fit_rad > 0:
            raise ArgumentError("fit_rad must be positive")
        img = np.asarray(image)
        if img.ndim != 2:
            raise ArgumentError("image must be a 2-D array")
        ny, nx = img.shape
        if inv_var is not None and np.shape(inv_var) != img.shape:
            raise ArgumentError("`inv_var` must be the same size as `image`")
        if model_rad == "auto" and inv_var is None:
            raise ArgumentError("model_rad=:auto requires inv_var (it needs a noise scale); pass inv_var or an "
                                "explicit scalar model_rad")
        model, mask = _sim_psf_and_mask(psf, fixed)
        self.y0, self.x0, self.f0 = _catalog(sources)
        n = len(self.y0)
        o = _abi.sim_opts_default()
        o.solver = {"lsqr": _abi.CP_SOLVER_LSQR, "lsmr": _abi.CP_SOLVER_LSMR, "cg": _abi.CP_SOLVER_CGLS}[solver]
        o.inner_iterations = inner_iterations
        o.linear_tol = linear_tol
        o.model_rad = 0.0 if model_rad == "auto" else float(model_rad)
        o.model_rad_max = float(model_rad_max)
        o.model_rad_nsigma = model_rad_nsigma
        o.max_step = max_step
        o.max_trials = max_trials
        o.max_iter = int(max_iter)
        o.x_tol, o.f_tol, o.g_tol = x_tol, f_tol, g_tol
        o.lambda_init, o.lambda_up, o.lambda_down = lambda_init, lambda_up, lambda_down
        o.lambda_min, o.lambda_max = lambda_min, lambda_max
        o.covariance = {None: _abi.CP_COV_DEFAULT, "known": _abi.CP_COV_KNOWN,
                        "reweighted": _abi.CP_COV_REWEIGHTED, "none": _abi.CP_COV_NONE}[covariance_estimator]
        o.free_mask = mask
        o.fixed_y = float(fixed.get("y", 0.0))
        o.fixed_x = float(fixed.get("x", 0.0))
        o.fixed_flux = float(fixed.get("flux", 0.0))
        o.fixed_bkg = float(fixed.get("bkg", 0.0))
        o.precision = {"mixed": _abi.CP_PREC_MIXED, "fp64": _abi.CP_PREC_FP64}[precision]
        o.tile = int(tile)
        o.layout = _abi.CP_LAYOUT_ROWMAJOR
        o.spread_model_fwhm = 0.0 if spread_model_fwhm is None else float(spread_model_fwhm)
        self.opts = o
        self.ny, self.nx, self.n = ny, nx, n
        self.fit_rad = float(fit_rad)
        self.show_trace = show_trace
        self.out = {k: np.zeros(n) for k in ("y", "x", "flux", "bkg", "y_err", "x_err", "flux_err", "bkg_err",
                                             "chisq", "qfit", "qfit_expected", "qfit_z", "crowding",
                                             "spread_model", "spread_model_err")}
        self.conv = np.zeros(n, np.uint8)
        self.valid = np.zeros(n, np.uint8)
        self.active = np.zeros(n, np.uint8)
        self.n_iter = np.zeros(n, np.int32)
        self.res = np.zeros(ny * nx, np.float32) if residual else None
        if show_trace and not trace:
            trace = int(max_iter) * int(max_trials)
        self.tr = np.zeros(max(int(trace), 1) * 8)
        s = _abi.SimResult()
        for k, v in self.out.items():
            setattr(s, k, _p(v, _abi.f64p))
        s.converged = _p(self.conv, _abi.u8p)
        s.valid = _p(self.valid, _abi.u8p)
        s.active = _p(self.active, _abi.u8p)
        s.n_iter = _p(self.n_iter, _abi.i32p)
        s.residual = _p(self.res, _abi.f32p)
        s.trace = _p(self.tr, _abi.f64p)
        s.trace_cap = int(trace)
        self.result = s
        self.desc = model.desc()
        self.im = _rowmajor32(img)
        self.iv = None if inv_var is None else _rowmajor32(inv_var)

    def run(self, ctx):
        rc = ctx.lib.cp_fit_simultaneous(ctx.handle, C.byref(self.desc), _p(self.im, _abi.f32p),
                                         _p(self.iv, _abi.f32p), self.ny, self.nx, _p(self.y0, _abi.f64p),
                                         _p(self.x0, _abi.f64p), _p(self.f0, _abi.f64p), self.n, self.fit_rad,
                                         C.byref(self.opts), C.byref(self.result))
        ctx.check(rc)

    def finish(self):
        n, s, tr = self.n, self.result, self.tr
        msgs = [f"star {i + 1}: excluded before the solve (no usable pixels or non-finite initial parameters)"
                for i in range(n) if not self.active[i]]
        resid = None if self.res is None else self.res.reshape(self.ny, self.nx)
        if self.show_trace:
            # the reference's per-trial line (psf_photometry_simultaneous.jl:1238-1246)
            na = max(int(s.n_active), 1)
            for row in tr[: 8 * s.trace_len].reshape(-1, 8):
                conv = math.floor(100 * (1 - row[5] / na) * 100) / 100
                print(f"Iter {int(row[0]):4d} trial {int(row[1]):2d} | converged = {conv:5.2f}% | "
                      f"cost = {row[2]:.4e} -> {row[3]:.4e} | \u03bb = {row[4]:.2e} | ||g|| = {row[7]:8.4f} | "
                      f"{'accepted' if row[3] < row[2] else 'rejected'}")
                if int(row[1]) == self.opts.max_trials and not row[3] < row[2]:
                    print(f"*** Max damping trials (`max_trials`) reached during outer iteration {int(row[0])} ***")
        out = self.out
        return MultiPassPhotResult(
            y=out["y"], x=out["x"], y_err=out["y_err"], x_err=out["x_err"], flux=out["flux"],
            flux_err=out["flux_err"], bkg=out["bkg"], bkg_err=out["bkg_err"], converged=self.conv.astype(bool),
            valid=self.valid.astype(bool), chisq=out["chisq"], qfit=out["qfit"], qfit_expected=out["qfit_expected"],
            qfit_z=out["qfit_z"], crowding=out["crowding"], spread_model=out["spread_model"],
            spread_model_err=out["spread_model_err"],
            n_iter=self.n_iter.astype(np.int64), n_passes=s.n_passes, n_failed=s.n_failed, failure_msgs=msgs,
            residual=resid,
            stats=dict(n_active=s.n_active, n_union=s.n_union, cost=s.cost_final, n_trials=s.n_trials_total,
                       n_inner=s.n_inner_total, trace=tr[: 8 * s.trace_len].reshape(-1, 8).copy()))


def fit_all_stars_simultaneous(image, psf, sources, fit_rad, *, ctx=None, **kw):
    """Simultaneous damped Gauss-Newton PSF photometry on the GPU, with the
    reference's keyword arguments and result (psf_photometry_simultaneous.jl:829-952):
    fixed, inv_var, solver, inner_iterations, linear_tol, model_rad, model_rad_max,
    model_rad_nsigma, max_step, max_trials, max_iter, x_tol, f_tol, g_tol, lambda_init,
    lambda_up, lambda_down, lambda_min, lambda_max, damping, show_trace,
    covariance_estimator, spread_model_fwhm.

    Build additions: ``precision`` ("mixed": float32 pixels/Jacobian/residuals
    with float64 reductions and star vectors; "fp64"), ``trace`` (rows of
    per-trial (iter, trial, cost, cost_cand, lambda, n_live, inner its) to keep),
    ``residual`` (False skips the residual image), ``ctx``.
    """
    call = _SimCall(image, psf, sources, fit_rad, **kw)
    call.run(ctx or default_context())
    return call.finish()


def fit_all_stars(image, psf, sources, fit_rad, *, n_passes=3, fixed=None, inv_var=None, spread_model_fwhm=None,
                  precision="mixed", residual=True, ctx=None, **kw):
    """Sequential multi-pass PSF photometry on the GPU (psf_photometry_single.jl:225-469):
    stars brightest first, each fitted with fit_star on the progressively
    subtracted residual over its +-fit_rad box, n_passes times.  Keyword
    arguments as the reference's (fixed, inv_var, spread_model_fwhm and
    fit_star's lm_irls keywords: lambda_init, ..., max_iter, x_tol, f_tol,
    g_tol, damping, covariance_estimator).  ``precision`` ("mixed": float32
    model and gradient with float64 sums and residual image; "fp64")."""
    if n_passes <= 0:
        raise ArgumentError(f"n_passes must be positive, got {n_passes}")
    if not isinstance(psf, (CircularGaussianPSF, CircularMoffatPSF, ImagePSF, GriddedPSFModel)):
        raise ArgumentError(f"unsupported PSF model {type(psf).__name__}")
    if not fit_rad > 0:
        raise ArgumentError("fit_rad must be positive")
    fixed = dict(fixed or {})
    img = np.asarray(image)
    if img.ndim != 2:
        raise ArgumentError("image must be a 2-D array")
    if inv_var is not None and np.shape(inv_var) != img.shape:
        raise ArgumentError("`inv_var` must be the same size as `image`")
    ny, nx = img.shape
    names = psf.props
    for k in fixed:
        if k not in names:
            raise ArgumentError(f"{type(psf).__name__} has no parameter {k!r}")
    mask = 0
    for k, nme in enumerate(names):
        if nme not in fixed:
            mask |= 1 << k
    if mask == 0:
        raise ArgumentError("all model parameters are fixed; nothing to fit")
    fixv = np.array([float(fixed[k]) if k in fixed else np.nan for k in names])
    # _extract_source_catalog / _build_params_matrix (:152-197): y, x, flux,
    # bkg from the catalogue, shape parameters from the template
    y, x, flux = _catalog(sources)
    get = (lambda k: sources[k]) if isinstance(sources, dict) else (lambda k: getattr(sources, k))
    has = (lambda k: k in sources) if isinstance(sources, dict) else (lambda k: hasattr(sources, k))
    bkg = np.ascontiguousarray(get("bkg"), np.float64) if has("bkg") else np.zeros_like(y)
    n = len(y)
    th0 = np.tile(psf.params(), (max(n, 1), 1))[:n].copy()
    cols = {"y": y, "x": x, "flux": flux, "bkg": bkg}
    for k, nme in enumerate(names):
        if nme in cols:
            th0[:, k] = cols[nme]
    th0 = np.ascontiguousarray(th0)
    out = {k: np.zeros(n) for k in ("y", "x", "flux", "bkg", "y_err", "x_err", "flux_err", "bkg_err", "chisq",
                                     "qfit", "qfit_expected", "qfit_z", "crowding", "spread_model",
                                     "spread_model_err")}
    conv = np.zeros(n, np.uint8)
    valid = np.zeros(n, np.uint8)
    nit = np.zeros(n, np.int32)
    res = np.zeros(ny * nx, np.float32) if residual else None
    r = _abi.SimResult()
    for k, v in out.items():
        setattr(r, k, _p(v, _abi.f64p))
    r.converged = _p(conv, _abi.u8p)
    r.valid = _p(valid, _abi.u8p)
    r.n_iter = _p(nit, _abi.i32p)
    r.residual = _p(res, _abi.f32p)
    o = _lm_opts(kw)
    im = _colmajor32(img)
    iv = None if inv_var is None else _colmajor32(inv_var)
    prec = {"mixed": _abi.CP_PREC_MIXED, "fp64": _abi.CP_PREC_FP64}[precision]
    ctx = ctx or default_context()
    rc = ctx.lib.cp_fit_all_stars(ctx.handle, C.byref(psf.desc()), _p(im, _abi.f32p), _p(iv, _abi.f32p), ny, nx,
                                  _p(th0, _abi.f64p), n, float(fit_rad), int(n_passes), mask, _p(fixv, _abi.f64p),
                                  C.byref(o), 0.0 if spread_model_fwhm is None else float(spread_model_fwhm), prec,
                                  C.byref(r))
    ctx.check(rc)
    if n == 0:
        resid = np.zeros((0, 0))
    else:
        resid = None if res is None else res.reshape(nx, ny).T
    msgs = []
    if r.n_failed:
        msgs = [f"{r.n_failed} / {n} star fits threw exceptions (too few valid inv_var pixels) and were marked "
                "invalid"]
    return MultiPassPhotResult(
        y=out["y"], x=out["x"], y_err=out["y_err"], x_err=out["x_err"], flux=out["flux"], flux_err=out["flux_err"],
        bkg=out["bkg"], bkg_err=out["bkg_err"], converged=conv.astype(bool), valid=valid.astype(bool),
        chisq=out["chisq"], qfit=out["qfit"], qfit_expected=out["qfit_expected"], qfit_z=out["qfit_z"],
        crowding=out["crowding"], spread_model=out["spread_model"], spread_model_err=out["spread_model_err"],
        n_iter=nit.astype(np.int64), n_passes=r.n_passes, n_failed=r.n_failed, failure_msgs=msgs, residual=resid)


# ---------------------------------------------------------------------------
# per-star LM
# ---------------------------------------------------------------------------
@dataclass
class LMResult:
    minimizer: np.ndarray
    minimum: float
    cost_init: float
    converged: bool
    x_converged: bool
    f_converged: bool
    g_converged: bool
    iterations: int
    lambda_final: float
    sigma_final: float
    cov: np.ndarray
    chisq: float


def _lm_opts(kw):
    o = _abi.lm_opts_default()
    names = {"λ_init": "lambda_init", "λ_up": "lambda_up", "λ_down": "lambda_down", "λ_min": "lambda_min",
             "λ_max": "lambda_max"}
    for k, v in kw.items():
        k = names.get(k, k)
        if k == "damping":
            v = {"marquardt": _abi.CP_DAMP_MARQUARDT, "levenberg": _abi.CP_DAMP_LEVENBERG,
                 "none": _abi.CP_DAMP_NONE}[v]
        elif k == "covariance_estimator":
            k = "covariance"
            v = {None: _abi.CP_COV_DEFAULT, "known": _abi.CP_COV_KNOWN,
                 "reweighted": _abi.CP_COV_REWEIGHTED, "none": _abi.CP_COV_NONE}[v]
        elif k in ("reweight", "scale_estimator", "weight_reset_tol", "show_trace"):
            if k == "reweight" and v is not None:
                raise ArgumentError("IRLS reweighting is not built on the GPU")
            continue
        if not hasattr(o, k):
            raise ArgumentError(f"unknown keyword {k!r}")
        setattr(o, k, v)
    return o


def fit_star_batch(model, cutouts, theta0=None, *, fixed=None, inv_var=None, ctx=None, **kw):
    """fit_star over many cutouts in one batched GPU launch.

    cutouts: list of 2-D arrays (each fit over all its pixels, coordinates
    1-based within the cutout).  theta0: (n, nparams) initial parameters in
    getproperties order (default: the model's own for every cutout).  fixed:
    names of the parameters held at their theta0 values.  Returns
    (theta [n, nparams], cov [n, p, p], stats structured list)."""
    if not isinstance(model, (CircularGaussianPSF, CircularMoffatPSF, ImagePSF, GriddedPSFModel)):
        raise ArgumentError(f"unsupported PSF model {type(model).__name__}")
    fixed = set(fixed or ())
    names = model.props
    npf = len(names)
    mask = 0
    for k, nme in enumerate(names):
        if nme not in fixed:
            mask |= 1 << k
    p = bin(mask).count("1")
    n = len(cutouts)
    ny = np.array([c.shape[0] for c in cutouts], np.int32)
    nx = np.array([c.shape[1] for c in cutouts], np.int32)
    off = np.zeros(n, np.int64)
    if n:
        off[1:] = np.cumsum(ny.astype(np.int64) * nx)[:-1]
    pix = np.concatenate([np.asarray(c, np.float32).T.ravel() for c in cutouts]) if n else np.zeros(1, np.float32)
    iv = None
    if inv_var is not None:
        iv = np.concatenate([np.asarray(v, np.float32).T.ravel() for v in inv_var])
    th0 = (np.tile(model.params(), (n, 1)) if theta0 is None else np.asarray(theta0, np.float64)).copy()
    th0 = np.ascontiguousarray(th0.reshape(n, npf))
    th = np.zeros_like(th0)
    cov = np.zeros((max(n, 1), p, p))
    st = (_abi.LMStat * max(n, 1))()
    o = _lm_opts(kw)
    ctx = ctx or default_context()
    rc = ctx.lib.cp_fit_star_batch(ctx.handle, C.byref(model.desc()), _p(pix, _abi.f32p), _p(off, _abi.i64p),
                                   _p(ny, _abi.i32p), _p(nx, _abi.i32p), _p(iv, _abi.f32p), _p(th0, _abi.f64p),
                                   mask, n, C.byref(o), _p(th, _abi.f64p), _p(cov, _abi.f64p), st)
    ctx.check(rc)
    return th, cov[:n], st


def fit_star(model, image, inds=None, *, fixed=None, inv_var=None, ctx=None, **kw):
    """fit_star(model, image, inds; fixed, inv_var, kws...) -> (best_model, LMResult)
    (psf_fitting.jl:229-308), as a one-cutout batch on the GPU.  ``inds`` is a
    (row_slice, col_slice) pair of 0-based Python slices; coordinates stay
    those of the full image."""
    fixed = dict(fixed or {})
    img = np.asarray(image, np.float64)
    if inv_var is not None and np.shape(inv_var) != img.shape:
        raise ArgumentError("`inv_var` must be the same size as `image`")
    r0, c0 = 0, 0
    if inds is not None:
        rs, cs = inds
        r0, c0 = rs.start or 0, cs.start or 0
        img = img[rs, cs]
        if inv_var is not None:
            inv_var = np.asarray(inv_var)[rs, cs]
    m = model.setproperties(**{k: v for k, v in fixed.items()}) if fixed else model
    th0 = m.params().copy()
    th0[0] -= r0
    th0[1] -= c0
    if isinstance(m, GriddedPSFModel) and (r0 or c0):
        # the cutout's coordinates are the image's shifted by (r0, c0): shift
        # the node grid with them, so the node blend at a star is unchanged
        # (each ImagePSF node is translation invariant)
        import copy
        m = copy.copy(m)
        m.grid_y = m.grid_y - r0
        m.grid_x = m.grid_x - c0
    if inv_var is not None:
        nvalid = int(np.sum(np.isfinite(inv_var) & (np.asarray(inv_var) > 0)))
        if nvalid <= len(model.props) - len(fixed):
            raise ArgumentError(f"too few valid inv_var pixels ({nvalid}) for "
                                f"{len(model.props) - len(fixed)} free parameters")
    th, cov, st = fit_star_batch(m, [img], th0[None, :], fixed=fixed.keys(),
                                 inv_var=None if inv_var is None else [inv_var], ctx=ctx, **kw)
    if st[0].status != 0:
        raise ArgumentError("fit_star rejected the cutout (too few observations for the free parameters)")
    best = th[0].copy()
    best[0] += r0
    best[1] += c0
    free = [k for k, nme in enumerate(model.props) if nme not in fixed]
    res = LMResult(minimizer=best[free], minimum=st[0].minimum, cost_init=st[0].cost_init,
                   converged=bool(st[0].converged), x_converged=bool(st[0].x_converged),
                   f_converged=bool(st[0].f_converged), g_converged=bool(st[0].g_converged),
                   iterations=st[0].iterations, lambda_final=st[0].lambda_final, sigma_final=float("nan"),
                   cov=cov[0], chisq=st[0].chisq)
    return model.setproperties(**dict(zip(model.props, best))), res


def evaluate_fg(model, py, px, ctx=None):
    """evaluate_fg at points (py, px) on the GPU: f [n], G [n, nparams]."""
    py = np.ascontiguousarray(np.atleast_1d(py), np.float64)
    px = np.ascontiguousarray(np.atleast_1d(px), np.float64)
    n = len(py)
    npf = len(model.props)
    f = np.zeros(n)
    G = np.zeros((n, npf))
    prm = np.ascontiguousarray(model.params())
    ctx = ctx or default_context()
    d = model.desc()
    ctx.check(ctx.lib.cp_evaluate_fg(ctx.handle, C.byref(d), _p(prm, _abi.f64p), _p(py, _abi.f64p),
                                     _p(px, _abi.f64p), n, _p(f, _abi.f64p), _p(G, _abi.f64p)))
    return f, G


def build_stamps(y0, x0, flux0, shape, fit_rad, w=None, ctx=None):
    """_build_stamps! + _touched_pixels on the GPU: (anchor_y, anchor_x,
    active (0-based, Morton order), pixels [n_active, S2], union_mask [ny, nx])."""
    ny, nx = shape
    y0 = np.ascontiguousarray(y0, np.float64)
    x0 = np.ascontiguousarray(x0, np.float64)
    flux0 = np.ascontiguousarray(flux0, np.float64)
    n = len(y0)
    R = int(np.ceil(fit_rad))
    S2 = (2 * R + 1) ** 2
    ay = np.zeros(max(n, 1), np.int32)
    ax = np.zeros(max(n, 1), np.int32)
    act = np.zeros(max(n, 1), np.int32)
    na = np.zeros(1, np.int32)
    pix = np.zeros(max(n, 1) * S2, np.int32)
    um = np.zeros(ny * nx, np.uint8)
    wv = None if w is None else _colmajor32(w)
    ctx = ctx or default_context()
    ctx.check(ctx.lib.cp_build_stamps(ctx.handle, _p(wv, _abi.f32p), ny, nx, _p(y0, _abi.f64p),
                                      _p(x0, _abi.f64p), _p(flux0, _abi.f64p), n, float(fit_rad),
                                      _p(ay, _abi.i32p), _p(ax, _abi.i32p), _p(act, _abi.i32p),
                                      _p(na, _abi.i32p), _p(pix, _abi.i32p), _p(um, _abi.u8p)))
    k = int(na[0])
    return ay[:n], ax[:n], act[:k], pix[: k * S2].reshape(k, S2), um.reshape(nx, ny).T
