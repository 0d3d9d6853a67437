"""Seeded synthetic inputs (simulate_image semantics, src/simulation.jl:211-498)
and the benchmark workloads built from them.

The random streams are the library's own (xoshiro256**), so the same seed
gives the same image on every machine; they are not Julia's StableRNG
streams, so reference fixtures are matched by construction, not by value.
"""
import math

import numpy as np

from . import _abi
from .psf import CircularGaussianPSF, CircularMoffatPSF, GriddedPSFModel, moffat_epsf

FLUX_DIST = {"uniform": 0, "loguniform": 1, "powerlaw": 2}


def _lib():
    return _abi.load_library()


def simulate_sources(seed, n, ylim, xlim, flux=(1e2, 1e5), flux_distribution="loguniform", flux_power=2.0):
    y, x, f = np.zeros(n), np.zeros(n), np.zeros(n)
    rc = _lib().cp_sim_sources(seed, n, ylim[0], ylim[1], xlim[0], xlim[1], FLUX_DIST[flux_distribution],
                               flux[0], flux[1], flux_power, y.ctypes.data_as(_abi.f64p),
                               x.ctypes.data_as(_abi.f64p), f.ctypes.data_as(_abi.f64p))
    if rc:
        raise ValueError("bad source distribution arguments")
    return y, x, f


def plummer_sources(seed, n, scale, centre, lo, hi, flux=(1e2, 1e6), flux_distribution="powerlaw",
                    flux_power=1.5):
    y, x, f = np.zeros(n), np.zeros(n), np.zeros(n)
    rc = _lib().cp_sim_sources_plummer(seed, n, scale, centre[0], centre[1], lo, hi, FLUX_DIST[flux_distribution],
                                       flux[0], flux[1], flux_power, y.ctypes.data_as(_abi.f64p),
                                       x.ctypes.data_as(_abi.f64p), f.ctypes.data_as(_abi.f64p))
    if rc:
        raise ValueError("bad Plummer arguments")
    return y, x, f


def render_sources(psf, shape, y, x, flux, model_radius):
    """Image [ny, nx] (float64) of the sources rendered over
    floor(y-r):ceil(y+r) x floor(x-r):ceil(x+r) (render_sources!)."""
    ny, nx = shape
    buf = np.zeros(ny * nx)
    y, x, flux = (np.ascontiguousarray(a, np.float64) for a in (y, x, flux))
    d = psf.desc()
    rc = _lib().cp_render_sources(d, ny, nx, y.ctypes.data_as(_abi.f64p), x.ctypes.data_as(_abi.f64p),
                                  flux.ctypes.data_as(_abi.f64p), len(y), float(model_radius),
                                  buf.ctypes.data_as(_abi.f64p))
    if rc:
        raise ValueError("render_sources: bad arguments")
    return buf.reshape(nx, ny).T.copy()


def add_noise(image, seed, gain=1.0, read_noise=0.0):
    """Poisson(gain * v) / gain then + N(0, read_noise^2) (add_noise!)."""
    a = np.ascontiguousarray(np.asarray(image, np.float64).T).ravel()
    rc = _lib().cp_add_noise(seed, a.ctypes.data_as(_abi.f64p), a.size, gain, read_noise)
    if rc:
        raise ValueError("add_noise: bad arguments")
    return a.reshape(image.shape[1], image.shape[0]).T.copy()


def _perturb(seed, y, x, f, pos_sigma=0.3, flux_lo=0.8, flux_hi=1.2):
    rng = np.random.default_rng(seed + 7919)
    return (y + rng.normal(0, pos_sigma, len(y)), x + rng.normal(0, pos_sigma, len(x)),
            f * rng.uniform(flux_lo, flux_hi, len(f)))


def config_gauss_small(seed=1, shape=(256, 256), n=200):
    """256x256, 200 stars, CircularGaussianPSF(fwhm=2.5), background 100,
    Poisson + N(0, 5^2), log-uniform fluxes 1e2-1e5, fit_rad 7 (15x15 stamp)."""
    bkg, rn = 100.0, 5.0
    y, x, f = simulate_sources(seed, n, (8.0, shape[0] - 8.0), (8.0, shape[1] - 8.0), (1e2, 1e5), "loguniform")
    psf = CircularGaussianPSF(fwhm=2.5)
    truth = render_sources(psf, shape, y, x, f, 7.0) + bkg
    img = add_noise(truth, seed, 1.0, rn) - bkg
    inv_var = np.full(shape, 1.0 / (bkg + rn * rn))
    y0, x0, f0 = _perturb(seed, y, x, f)
    return dict(image=img, inv_var=inv_var, psf=psf, fixed={"fwhm": 2.5, "bkg": 0.0},
                sources={"y": y0, "x": x0, "flux": f0}, truth=(y, x, f), fit_rad=7, model_rad=7.0,
                max_iter=40)


def epsf_c2():
    """ePSF of the crowded-field workloads: CircularMoffat beta=2.5, FWHM 2.2,
    101x101 grid at 4x oversampling with origin (51, 51), pixel-response
    convolved, normalised to sum 16, as a one-node GriddedPSFModel."""
    node = moffat_epsf(fwhm=2.2, beta=2.5, size=101, oversampling=4, origin=(51.0, 51.0))
    return GriddedPSFModel([node], [0.0], [0.0])


def config_epsf_crowded(seed=2, shape=(4096, 4096), n=200_000, max_iter=10):
    """4096x4096 FP32 image, background 50 + 0.01 (x + y) subtracted exactly,
    Poisson + N(0, 4^2), n stars uniform in [8, side-8], power-law fluxes
    (slope -1.5) over 1e2-1e6, ePSF 101x101 at 4x, fit_rad 3 (7x7 stamps),
    model_rad 12, 10 outer iterations."""
    ny, nx = shape
    rn = 4.0
    y, x, f = simulate_sources(seed, n, (8.0, ny - 8.0), (8.0, nx - 8.0), (1e2, 1e6), "powerlaw", 1.5)
    psf = epsf_c2()
    gy = np.arange(1, ny + 1)[:, None]
    gx = np.arange(1, nx + 1)[None, :]
    bkg = 50.0 + 0.01 * (gx + gy)
    truth = render_sources(psf, shape, y, x, f, 12.0) + bkg
    img = (add_noise(truth, seed, 1.0, rn) - bkg).astype(np.float32)
    inv_var = (1.0 / (bkg + rn * rn)).astype(np.float32)
    y0, x0, f0 = _perturb(seed, y, x, f)
    return dict(image=img, inv_var=inv_var, psf=psf, fixed={"bkg": 0.0},
                sources={"y": y0, "x": x0, "flux": f0}, truth=(y, x, f), fit_rad=3, model_rad=12.0,
                max_iter=max_iter)


def serving_tile(seed, side, model="epsf", max_iter=10):
    """One request of the serving workload: a side x side tile with
    min(200000, max(1000, side^2 / 80)) stars, background 50 subtracted
    exactly, Poisson + N(0, 4^2), power-law fluxes 1e2-1e6; model "epsf"
    (the crowded-field ePSF, fit_rad 3, model_rad 12) or "gauss"
    (CircularGaussianPSF fwhm 2.5, fit_rad 3, model_rad 7)."""
    n = int(min(200_000, max(1000, side * side // 80)))
    bkg, rn = 50.0, 4.0
    y, x, f = simulate_sources(seed, n, (8.0, side - 8.0), (8.0, side - 8.0), (1e2, 1e6), "powerlaw", 1.5)
    if model == "epsf":
        psf, fixed, mr = epsf_c2(), {"bkg": 0.0}, 12.0
    else:
        psf, fixed, mr = CircularGaussianPSF(fwhm=2.5), {"fwhm": 2.5, "bkg": 0.0}, 7.0
    truth = render_sources(psf, (side, side), y, x, f, mr) + bkg
    img = (add_noise(truth, seed, 1.0, rn) - bkg).astype(np.float32)
    inv_var = np.full((side, side), 1.0 / (bkg + rn * rn), np.float32)
    y0, x0, f0 = _perturb(seed, y, x, f)
    return dict(image=img, inv_var=inv_var, psf=psf, fixed=fixed, sources={"y": y0, "x": x0, "flux": f0},
                truth=(y, x, f), fit_rad=3, model_rad=mr, max_iter=max_iter, n=n, side=side, model=model)


def config_plummer(seed=4, shape=(8192, 8192), n=1_000_000, scale=1500.0, max_iter=10):
    """As the crowded ePSF field at 8192^2 with a Plummer-clustered population."""
    ny, nx = shape
    rn = 4.0
    y, x, f = plummer_sources(seed, n, scale, ((ny + 1) / 2, (nx + 1) / 2), 8.0, ny - 8.0)
    psf = epsf_c2()
    gy = np.arange(1, ny + 1)[:, None]
    gx = np.arange(1, nx + 1)[None, :]
    bkg = 50.0 + 0.01 * (gx + gy)
    truth = render_sources(psf, shape, y, x, f, 12.0) + bkg
    img = (add_noise(truth, seed, 1.0, rn) - bkg).astype(np.float32)
    inv_var = (1.0 / (bkg + rn * rn)).astype(np.float32)
    y0, x0, f0 = _perturb(seed, y, x, f)
    return dict(image=img, inv_var=inv_var, psf=psf, fixed={"bkg": 0.0},
                sources={"y": y0, "x": x0, "flux": f0}, truth=(y, x, f), fit_rad=3, model_rad=12.0,
                max_iter=max_iter)


def config_plummer_crop(seed=4, side=512, offset=(0.0, 0.0), n_full=1_000_000, full=8192, scale=1500.0,
                        max_iter=10):
    """A side x side crop of the Plummer-clustered 8192^2 field (1e6 stars,
    scale radius 1500 px) at the same local density: the crop's lower corner
    sits at the field centre minus side / 2, shifted by `offset` pixels.  Stars
    of the full population whose light reaches the crop are rendered into it;
    the catalogue holds those inside [8, side - 8] of the crop (crop
    coordinates), perturbed like every other workload."""
    rn = 4.0
    c = (full + 1) / 2
    y, x, f = plummer_sources(seed, n_full, scale, (c, c), 8.0, full - 8.0)
    y0c = math.floor(c - side / 2 + offset[0])
    x0c = math.floor(c - side / 2 + offset[1])
    yy, xx = y - y0c, x - x0c
    near = (yy > -13) & (yy < side + 13) & (xx > -13) & (xx < side + 13)
    yy, xx, ff = yy[near], xx[near], f[near]
    psf = epsf_c2()
    gy = np.arange(1, side + 1)[:, None] + y0c
    gx = np.arange(1, side + 1)[None, :] + x0c
    bkg = 50.0 + 0.01 * (gx + gy)
    truth = render_sources(psf, (side, side), yy, xx, ff, 12.0) + bkg
    img = (add_noise(truth, seed, 1.0, rn) - bkg).astype(np.float32)
    inv_var = (1.0 / (bkg + rn * rn)).astype(np.float32)
    inside = (yy >= 8) & (yy <= side - 8) & (xx >= 8) & (xx <= side - 8)
    ty, tx, tf = yy[inside], xx[inside], ff[inside]
    py, px, pf = _perturb(seed, ty, tx, tf)
    return dict(image=img, inv_var=inv_var, psf=psf, fixed={"bkg": 0.0},
                sources={"y": py, "x": px, "flux": pf}, truth=(ty, tx, tf), fit_rad=3, model_rad=12.0,
                max_iter=max_iter, n_rendered=int(near.sum()))


def config_cutouts(seed=3, n=1_000_000, beta=2.5):
    """Isolated-star cutouts with odd sides uniform over 9..31, CircularMoffat
    (FWHM ~ U(1.8, 3.0), beta fixed), flux log-uniform 1e2-1e5, background
    U(10, 200), Poisson noise.  Returns packed arrays for fit_star_batch:
    pix (float32), offsets, ny, nx, theta0 (n, 6), truth (n, 6), free mask."""
    rng = np.random.default_rng(seed)
    sides = rng.choice(np.arange(9, 32, 2), n)
    fwhm = rng.uniform(1.8, 3.0, n)
    alpha = fwhm / (2 * np.sqrt(2 ** (1 / beta) - 1))
    flux = np.exp(np.log(1e2) + rng.random(n) * (np.log(1e5) - np.log(1e2)))
    bkg = rng.uniform(10, 200, n)
    cy = (sides + 1) / 2 + rng.uniform(-0.5, 0.5, n)
    cx = (sides + 1) / 2 + rng.uniform(-0.5, 0.5, n)
    npx = sides.astype(np.int64) ** 2
    off = np.zeros(n, np.int64)
    off[1:] = np.cumsum(npx)[:-1]
    pix = np.zeros(int(npx.sum()), np.float32)
    for s in np.unique(sides):
        idx = np.nonzero(sides == s)[0]
        ii = np.arange(1, s + 1, dtype=np.float64)
        dy = ii[None, :, None] - cy[idx, None, None]     # [k, row i, 1]
        dx = ii[None, None, :] - cx[idx, None, None]     # [k, 1, col j]
        a2 = (alpha[idx] ** 2)[:, None, None]
        amp = (flux[idx] * (beta - 1) / (np.pi * alpha[idx] ** 2))[:, None, None]
        model = amp * (1 + (dx * dx + dy * dy) / a2) ** (-beta) + bkg[idx, None, None]
        noisy = rng.poisson(model).astype(np.float32)
        # column-major per cutout: transpose (row, col) -> (col, row)
        flat = np.ascontiguousarray(noisy.transpose(0, 2, 1)).reshape(len(idx), -1)
        pos = off[idx][:, None] + np.arange(s * s)[None, :]
        pix[pos.ravel()] = flat.ravel()
    truth = np.stack([cy, cx, alpha, np.full(n, beta), flux, bkg], 1)
    th0 = truth.copy()
    th0[:, 0] += rng.normal(0, 0.3, n)
    th0[:, 1] += rng.normal(0, 0.3, n)
    th0[:, 4] *= rng.uniform(0.8, 1.2, n)
    th0[:, 5] *= rng.uniform(0.9, 1.1, n)
    # free (y, x, flux, bkg) = slots 0, 1, 4, 5
    return dict(pix=pix, offsets=off, ny=sides.astype(np.int32), nx=sides.astype(np.int32),
                theta0=np.ascontiguousarray(th0), truth=truth, free_mask=0b110011, beta=beta)
