"""PSF model types with the reference's names, fields and parameter order.

Mirrors cgarling/CrowdPhot.jl:
  CircularGaussianPSF  src/psf/parametric_models.jl:62-113   (y, x, fwhm, flux, bkg)
  CircularMoffatPSF    src/psf/parametric_models.jl:672-724  (y, x, alpha, beta, flux, bkg)
  ImagePSF             src/psf/empirical_models.jl:76-157    (y, x, flux, bkg)
  GriddedPSFModel      src/psf/gridded_psf.jl:51-144         (y, x, flux, bkg)

The classes are parameter containers; evaluation and fitting run on the GPU
through the C ABI (``desc()`` builds the cp_psf_desc the library consumes).
"""
import ctypes as C
import math

import numpy as np

from . import _abi

GAUSS_PRE = -4.0 * math.log(2.0)  # src/psf/parametric_models.jl:18


class _Model:
    props = ()
    kind = None

    def getproperties(self):
        return {k: getattr(self, k) for k in self.props}

    def setproperties(self, **patch):
        """Copy with the named parameters replaced (ConstructionBase.setproperties)."""
        import copy
        m = copy.copy(self)
        for k, v in patch.items():
            if k not in self.props:
                raise ValueError(f"{type(self).__name__} has no parameter {k!r}")
            setattr(m, k, float(v))
        return m

    def params(self):
        return np.array([getattr(self, k) for k in self.props], dtype=np.float64)

    def __repr__(self):
        inner = ", ".join(f"{k}={getattr(self, k)!r}" for k in self.props)
        return f"{type(self).__name__}({inner})"


class CircularGaussianPSF(_Model):
    props = ("y", "x", "fwhm", "flux", "bkg")
    kind = _abi.CP_PSF_CIRCULAR_GAUSSIAN

    def __init__(self, y=0.0, x=0.0, fwhm=1.0, flux=1.0, bkg=0.0):
        self.y, self.x, self.fwhm, self.flux, self.bkg = map(float, (y, x, fwhm, flux, bkg))

    def desc(self):
        d = _abi.PsfDesc()
        d.kind = self.kind
        d.fwhm = self.fwhm
        return d


class CircularMoffatPSF(_Model):
    props = ("y", "x", "alpha", "beta", "flux", "bkg")
    kind = _abi.CP_PSF_CIRCULAR_MOFFAT

    def __init__(self, y=0.0, x=0.0, alpha=1.0, beta=2.5, flux=1.0, bkg=0.0):
        self.y, self.x, self.alpha, self.beta, self.flux, self.bkg = map(
            float, (y, x, alpha, beta, flux, bkg))

    @staticmethod
    def alpha_from_fwhm(fwhm, beta):
        """alpha with FWHM = 2 alpha sqrt(2^(1/beta) - 1) (parametric_models.jl:689-692)."""
        return fwhm / (2.0 * math.sqrt(2.0 ** (1.0 / beta) - 1.0))

    def desc(self):
        d = _abi.PsfDesc()
        d.kind = self.kind
        d.alpha = self.alpha
        d.beta = self.beta
        return d


class ImagePSF(_Model):
    """Oversampled ePSF on a grid, evaluated with Keys (a = -1/2) bicubic
    interpolation (empirical_models.jl:239-318).  ``data`` is indexed
    [row = y, column = x]; ``oversampling`` is (x factor, y factor)."""
    props = ("y", "x", "flux", "bkg")
    kind = _abi.CP_PSF_GRIDDED_IMAGE

    def __init__(self, data, y=0.0, x=0.0, flux=1.0, bkg=0.0, origin=None, oversampling=1,
                 fill_value=0.0, normalize=False):
        data = np.array(data, dtype=np.float64)
        if data.ndim != 2:
            raise ValueError("`data` must be a 2D matrix")
        if min(data.shape) < 4:
            raise ValueError("each axis of `data` must have at least 4 samples")
        if not np.all(np.isfinite(data)):
            raise ValueError("all elements of `data` must be finite")
        if np.isscalar(oversampling):
            os_ = (int(oversampling), int(oversampling))
        else:
            os_ = tuple(int(v) for v in oversampling)
        if normalize:
            s = data.sum()
            if not (np.isfinite(s) and s > 0):
                raise ValueError("cannot normalize PSF data with non-positive sum")
            data = data * (os_[0] * os_[1]) / s
        if origin is None:
            origin = ((data.shape[0] + 1) / 2.0, (data.shape[1] + 1) / 2.0)
        elif isinstance(origin, dict):
            origin = (float(origin["y"]), float(origin["x"]))
        self.data = data
        self.origin = (float(origin[0]), float(origin[1]))  # (y, x)
        self.oversampling = os_
        self.fill_value = float(fill_value)
        self.y, self.x, self.flux, self.bkg = map(float, (y, x, flux, bkg))

    def extent(self):
        """Detector-pixel bounds of the grid (empirical_models.jl:167-177)."""
        ny, nx = self.data.shape
        sx, sy = self.oversampling
        oy, ox = self.origin
        return ((self.y - (oy - 1) / sy, self.y + (ny - oy) / sy),
                (self.x - (ox - 1) / sx, self.x + (nx - ox) / sx))

    def desc(self):
        return GriddedPSFModel([self], [0.0], [0.0]).desc()


class GriddedPSFModel(_Model):
    """Bilinear blend of ImagePSF nodes on a rectangular (y, x) grid
    (gridded_psf.jl:51-271).  All nodes must share one grid geometry."""
    props = ("y", "x", "flux", "bkg")
    kind = _abi.CP_PSF_GRIDDED_IMAGE

    def __init__(self, psfs, grid_y, grid_x, y=0.0, x=0.0, flux=1.0, bkg=0.0):
        if len(psfs) < 1:
            raise ValueError("`psfs` must be non-empty")
        if not (len(grid_y) == len(grid_x) == len(psfs)):
            raise ValueError("`psfs`, `grid_y`, and `grid_x` must all have the same length")
        n0 = psfs[0]
        for p in psfs:
            if not isinstance(p, ImagePSF):
                raise ValueError("GriddedPSFModel nodes must be ImagePSF in this build")
            if (p.data.shape != n0.data.shape or p.oversampling != n0.oversampling
                    or p.origin != n0.origin or p.fill_value != n0.fill_value):
                raise ValueError("all ImagePSF nodes must share shape, oversampling, origin, fill")
        gy = np.asarray(grid_y, dtype=np.float64)
        gx = np.asarray(grid_x, dtype=np.float64)
        if not (np.all(np.isfinite(gy)) and np.all(np.isfinite(gx))):
            raise ValueError("`grid_y` and `grid_x` must be finite")
        if len(psfs) > 1:
            uy, ux = np.unique(gy), np.unique(gx)
            if len(uy) < 2 or len(ux) < 2:
                raise ValueError("`grid_y` and `grid_x` must each contain at least 2 unique values")
            if len(uy) * len(ux) != len(psfs):
                raise ValueError("node positions do not form a complete rectangular grid")
            if len(set(zip(gy.tolist(), gx.tolist()))) != len(psfs):
                raise ValueError("duplicate node position")
        self.psfs = list(psfs)
        self.grid_y, self.grid_x = gy, gx
        self.y, self.x, self.flux, self.bkg = map(float, (y, x, flux, bkg))

    def desc(self):
        n0 = self.psfs[0]
        # node grids column-major (y fastest) as float32
        stack = np.stack([np.asfortranarray(p.data, dtype=np.float32) for p in self.psfs])
        self._grids = np.ascontiguousarray(
            np.concatenate([s.ravel(order="F") for s in stack]), dtype=np.float32)
        self._gy = np.ascontiguousarray(self.grid_y, dtype=np.float64)
        self._gx = np.ascontiguousarray(self.grid_x, dtype=np.float64)
        d = _abi.PsfDesc()
        d.kind = self.kind
        d.n_nodes = len(self.psfs)
        d.gny, d.gnx = n0.data.shape
        d.grids = self._grids.ctypes.data_as(_abi.f32p)
        d.os_x, d.os_y = n0.oversampling
        d.origin_y, d.origin_x = n0.origin
        d.fill_value = n0.fill_value
        d.node_y = self._gy.ctypes.data_as(_abi.f64p)
        d.node_x = self._gx.ctypes.data_as(_abi.f64p)
        d._keep = (self._grids, self._gy, self._gx)
        return d


def pixel_response_kernel(n):
    """Box2DKernel(width=n) pixel-response kernel (src/psf/PSF.jl:512-518)."""
    sz = n if n % 2 == 1 else n + 1
    half = n / 2.0
    centers = np.arange(sz) - (sz - 1) / 2.0
    w1 = np.clip(np.minimum(centers + 0.5, half) - np.maximum(centers - 0.5, -half), 0, None) / n
    return np.outer(w1, w1)


def moffat_epsf(fwhm=2.2, beta=2.5, size=101, oversampling=4, origin=None):
    """ePSF grid of a circular Moffat: point-sampled on a size x size grid at
    `oversampling`, convolved with pixel_response_kernel(oversampling) (zero
    padding) and normalised to sum = oversampling^2 (empirical_models.jl:91-140)."""
    alpha = CircularMoffatPSF.alpha_from_fwhm(fwhm, beta)
    if origin is None:
        origin = ((size + 1) / 2.0, (size + 1) / 2.0)
    i = (np.arange(1, size + 1) - origin[0]) / oversampling
    j = (np.arange(1, size + 1) - origin[1]) / oversampling
    r2 = i[:, None] ** 2 + j[None, :] ** 2
    grid = (beta - 1) / (math.pi * alpha * alpha) * (1 + r2 / (alpha * alpha)) ** (-beta)
    k = pixel_response_kernel(oversampling)
    h = k.shape[0] // 2
    pad = np.pad(grid, h)
    out = np.zeros_like(grid)
    for a in range(k.shape[0]):
        for b in range(k.shape[1]):
            out += k[a, b] * pad[a:a + size, b:b + size]
    return ImagePSF(out, oversampling=oversampling, origin=origin, normalize=True)
