"""Host mirror of CrowdPhot.jl's detection step over libcrowdphot_hip.so.

  correlate / correlate!   src/correlation.jl:108-140
  findlocalmaxima          src/detection.jl:240-324
  matched_filter           src/detection.jl:87-232
  MatchedFilterResult      src/detection.jl:8-48
  render                   src/psf/PSF.jl:320-350 (the kernel a PSF model gives)

Images are numpy arrays indexed [row = y, column = x]; peaks are 1-based
(row, column) pairs, as the reference's CartesianIndex.  The correlations,
the significance map and the peak search run on the GPU; only the tiny
kernel set-up (normalisation, PRF sampling) is host work.
"""
import ctypes as C
import math
from dataclasses import dataclass
from typing import List, Optional, Tuple

import numpy as np

from . import _abi
from .photometry import ArgumentError, _p, default_context
from .psf import CircularGaussianPSF, CircularMoffatPSF, GriddedPSFModel, ImagePSF

_BORDER = {"replicate": _abi.CP_BORDER_REPLICATE, "zero": _abi.CP_BORDER_ZERO}


def _border(b):
    b = b.lstrip(":") if isinstance(b, str) else b
    if b not in _BORDER:
        raise ArgumentError(f"unknown border mode :{b}; use :replicate or :zero")
    return _BORDER[b]


def _kernel(kernel):
    k = np.asarray(kernel, dtype=np.float64)
    if k.ndim != 2:
        raise ArgumentError("kernel must be a matrix")
    m, n = k.shape
    if m % 2 == 0 or n % 2 == 0:
        raise ArgumentError(f"kernel dimensions must be odd (got {m}x{n}); the correlation center is at "
                            "((m+1)/2, (n+1)/2), which requires odd-sized kernels")
    return np.asfortranarray(k).ravel(order="F"), m, n


def _col64(img):
    a = np.asarray(img, dtype=np.float64)
    if a.ndim != 2:
        raise ArgumentError("image must be a 2-D array")
    return np.ascontiguousarray(a.T).ravel(), a.shape


def correlate(img, kernel, border="replicate", ctx=None):
    """2-D correlation (no kernel flip) with the reference's centre
    ((m+1)/2, (n+1)/2) and border handling; float64 result of img's shape."""
    kv, kr, kc = _kernel(kernel)
    a, (ny, nx) = _col64(img)
    out = np.zeros(ny * nx)
    ctx = ctx or default_context()
    ctx.check(ctx.lib.cp_correlate(ctx.handle, _p(a, _abi.f64p), ny, nx, _p(kv, _abi.f64p), kr, kc,
                                   _border(border), _p(out, _abi.f64p)))
    return out.reshape(nx, ny).T


def findlocalmaxima(img, edges=True, ctx=None) -> List[Tuple[int, int]]:
    """1-based (row, column) of every pixel strictly above its 8 neighbours,
    in the reference's order."""
    a, (ny, nx) = _col64(img)
    cap = max(1, ((ny + 1) // 2 + 1) * ((nx + 1) // 2 + 1))
    py = np.zeros(cap, np.int32)
    px = np.zeros(cap, np.int32)
    n = C.c_int32(0)
    ctx = ctx or default_context()
    ctx.check(ctx.lib.cp_findlocalmaxima(ctx.handle, _p(a, _abi.f64p), ny, nx, int(bool(edges)), cap,
                                         _p(py, _abi.i32p), _p(px, _abi.i32p), C.byref(n)))
    return list(zip(py[: n.value].tolist(), px[: n.value].tolist()))


# --------------------------------------------------------------------------
# kernels from models
# --------------------------------------------------------------------------
class CircularGaussianPRF:
    """Pixel-integrated circular Gaussian (parametric_models.jl:445-470), used
    here only to build detection kernels."""

    def __init__(self, fwhm, y=0.0, x=0.0, flux=1.0, bkg=0.0):
        self.fwhm, self.y, self.x, self.flux, self.bkg = map(float, (fwhm, y, x, flux, bkg))

    def extent_half(self):
        h = 5 * self.fwhm / 2
        return h, h

    def evaluate(self, py, px):
        a = 2 * math.sqrt(math.log(2)) / self.fwhm
        ex = math.erf(a * (px - self.x + 0.5)) - math.erf(a * (px - self.x - 0.5))
        ey = math.erf(a * (py - self.y + 0.5)) - math.erf(a * (py - self.y - 0.5))
        return self.flux / 4 * ex * ey + self.bkg


class GaussianPRF:
    """Pixel-integrated elliptical Gaussian with theta = 0
    (parametric_models.jl:550-590), for detection kernels."""

    def __init__(self, y_fwhm, x_fwhm, y=0.0, x=0.0, flux=1.0, bkg=0.0):
        self.y_fwhm, self.x_fwhm = float(y_fwhm), float(x_fwhm)
        self.y, self.x, self.flux, self.bkg = map(float, (y, x, flux, bkg))

    def extent_half(self):
        return 5 * self.y_fwhm / 2, 5 * self.x_fwhm / 2

    def evaluate(self, py, px):
        c = 2 * math.sqrt(math.log(2))
        ax, ay = c / self.x_fwhm, c / self.y_fwhm
        ex = math.erf(ax * (px - self.x + 0.5)) - math.erf(ax * (px - self.x - 0.5))
        ey = math.erf(ay * (py - self.y + 0.5)) - math.erf(ay * (py - self.y - 0.5))
        return self.flux / 4 * ex * ey + self.bkg


def render(model, ctx=None):
    """Model evaluated on the odd grid centred on its rounded centroid that
    covers its extent (PSF.jl:320-350): 5 FWHM for the analytic models, the
    grid footprint for ImagePSF / GriddedPSFModel."""
    y0, x0 = model.y, model.x
    if isinstance(model, (CircularGaussianPRF, GaussianPRF)):
        dy, dx = model.extent_half()
        ylo, yhi, xlo, xhi = y0 - dy, y0 + dy, x0 - dx, x0 + dx
    elif isinstance(model, CircularGaussianPSF):
        h = 5 * model.fwhm / 2
        ylo, yhi, xlo, xhi = y0 - h, y0 + h, x0 - h, x0 + h
    elif isinstance(model, CircularMoffatPSF):
        h = 5 * (2 * model.alpha * math.sqrt(2 ** (1 / model.beta) - 1)) / 2
        ylo, yhi, xlo, xhi = y0 - h, y0 + h, x0 - h, x0 + h
    elif isinstance(model, ImagePSF):
        (ylo, yhi), (xlo, xhi) = model.extent()
    elif isinstance(model, GriddedPSFModel):
        # every node shares the grid geometry, so the union of the corner
        # nodes' extents is one node's footprint
        (ylo, yhi), (xlo, xhi) = model.psfs[0].setproperties(y=y0, x=x0).extent()
    else:
        raise ArgumentError(f"cannot render {type(model).__name__}")
    hx = max(math.ceil(x0 - xlo), math.ceil(xhi - x0))
    hy = max(math.ceil(y0 - ylo), math.ceil(yhi - y0))
    xc, yc = int(np.rint(x0)), int(np.rint(y0))
    ys = yc - hy + np.arange(2 * hy + 1, dtype=np.float64)
    xs = xc - hx + np.arange(2 * hx + 1, dtype=np.float64)
    if isinstance(model, (CircularGaussianPRF, GaussianPRF)):
        return np.array([[model.evaluate(y, x) for x in xs] for y in ys])
    from .photometry import evaluate_fg
    py = np.repeat(ys, len(xs))
    px = np.tile(xs, len(ys))
    f, _ = evaluate_fg(model, py, px, ctx=ctx)
    return f.reshape(len(ys), len(xs))


@dataclass
class MatchedFilterResult:
    image: np.ndarray
    inv_var: Optional[np.ndarray]
    smoothed_image: np.ndarray
    smoothed_inv_var: Optional[np.ndarray]
    significance_map: np.ndarray
    kernel: np.ndarray
    kernel_norm: float
    peaks: List[Tuple[int, int]]
    peak_significances: np.ndarray
    peak_fluxes: np.ndarray


def matched_filter(image, kernel, *, inv_var=None, normalize_zerosum=True, sigma=5.0, border="replicate",
                   ctx=None):
    """Matched-filter detection (detection.jl:87-232).  ``kernel`` is a
    matrix, a PSF model (rendered), a number (CircularGaussianPRF FWHM) or a
    (y_fwhm, x_fwhm) pair (GaussianPRF)."""
    ctx = ctx or default_context()
    if isinstance(kernel, (int, float)) and not isinstance(kernel, bool):
        kernel = render(CircularGaussianPRF(kernel), ctx)
    elif isinstance(kernel, tuple) and len(kernel) == 2 and all(np.isscalar(v) for v in kernel):
        kernel = render(GaussianPRF(kernel[0], kernel[1]), ctx)
    elif not isinstance(kernel, np.ndarray) and hasattr(kernel, "props"):
        kernel = render(kernel, ctx)
    kv, kr, kc = _kernel(kernel)
    img = np.asarray(image)
    if img.ndim != 2:
        raise ArgumentError("image must be a 2-D array")
    ny, nx = img.shape
    if inv_var is not None and np.shape(inv_var) != img.shape:
        raise ArgumentError(f"inv_var size {np.shape(inv_var)} must match image size {img.shape}")
    im = np.ascontiguousarray(np.asarray(img, np.float32).T).ravel()
    iv = None
    ivc = None
    if inv_var is not None:
        ivc = np.asarray(inv_var, np.float64)
        ivc = np.where(np.isfinite(ivc), ivc, 0.0)
        iv = np.ascontiguousarray(ivc.astype(np.float32).T).ravel()
    P = ny * nx
    sm = np.zeros(P)
    sig = np.zeros(P)
    den = np.zeros(P) if inv_var is not None else None
    kout = np.zeros(kr * kc)
    cap = max(1, ((ny + 1) // 2 + 1) * ((nx + 1) // 2 + 1))
    py = np.zeros(cap, np.int32)
    px = np.zeros(cap, np.int32)
    ps = np.zeros(cap)
    pf = np.zeros(cap)
    r = _abi.MFResult()
    r.smoothed = _p(sm, _abi.f64p)
    r.smoothed_inv_var = _p(den, _abi.f64p)
    r.significance = _p(sig, _abi.f64p)
    r.kernel = _p(kout, _abi.f64p)
    r.peaks_cap = cap
    r.peak_y, r.peak_x = _p(py, _abi.i32p), _p(px, _abi.i32p)
    r.peak_significance, r.peak_flux = _p(ps, _abi.f64p), _p(pf, _abi.f64p)
    ctx.check(ctx.lib.cp_matched_filter(ctx.handle, _p(im, _abi.f32p), _p(iv, _abi.f32p), ny, nx,
                                        _p(kv, _abi.f64p), kr, kc, int(bool(normalize_zerosum)), float(sigma),
                                        _border(border), C.byref(r)))
    n = r.n_peaks
    shape = lambda v: None if v is None else v.reshape(nx, ny).T  # noqa: E731
    return MatchedFilterResult(
        image=image, inv_var=ivc, smoothed_image=shape(sm), smoothed_inv_var=shape(den),
        significance_map=shape(sig), kernel=kout.reshape(kc, kr).T, kernel_norm=r.kernel_norm,
        peaks=list(zip(py[:n].tolist(), px[:n].tolist())), peak_significances=ps[:n].copy(),
        peak_fluxes=pf[:n].copy())
