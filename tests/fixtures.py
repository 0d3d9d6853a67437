"""Small seeded synthetic inputs for CPU tests (numpy; closed-form Gaussian render)."""
import math

import numpy as np

GAUSS_PRE = -4.0 * math.log(2.0)


def render_gauss(shape, ys, xs, fs, fwhm, radius):
    """image[ny, nx] (row = y) with each circular Gaussian summed over
    floor(y-r):ceil(y+r) x floor(x-r):ceil(x+r) (render_sources!, simulation.jl:321-347)."""
    ny, nx = shape
    img = np.zeros(shape)
    norm = -(math.pi * fwhm * fwhm / GAUSS_PRE)
    for y, x, f in zip(ys, xs, fs):
        y0, y1 = max(1, math.floor(y - radius)), min(ny, math.ceil(y + radius))
        x0, x1 = max(1, math.floor(x - radius)), min(nx, math.ceil(x + radius))
        gy = np.arange(y0, y1 + 1)[:, None]
        gx = np.arange(x0, x1 + 1)[None, :]
        sq = ((gx - x) ** 2 + (gy - y) ** 2) / (fwhm * fwhm)
        img[y0 - 1:y1, x0 - 1:x1] += (f / norm) * np.exp(GAUSS_PRE * sq)
    return img


def separated_positions(rng, n, shape, border, min_sep):
    ys, xs = [], []
    while len(ys) < n:
        y = rng.uniform(border, shape[0] - border)
        x = rng.uniform(border, shape[1] - border)
        if all((y - a) ** 2 + (x - b) ** 2 >= min_sep ** 2 for a, b in zip(ys, xs)):
            ys.append(y)
            xs.append(x)
    return np.array(ys), np.array(xs)


def crowded_field(seed, shape=(150, 150), n=200, fwhm=2.0, bkg=50.0, read_noise=2.0,
                  flux=(100.0, 3000.0), radius=6, border=8, min_sep=0.0):
    """Crowded noisy field in the manner of the reference crowded fixture
    (test/photometry/psf_photometry_simultaneous.jl:255-270): Poisson + Gaussian
    noise, power-law fluxes (slope 2), background subtracted; candidate centres
    closer than min_sep to an accepted one are redrawn (simulate_sources,
    src/simulation.jl:240-273)."""
    rng = np.random.default_rng(seed)
    if min_sep > 0:
        ys, xs = np.zeros(n), np.zeros(n)
        k = 0
        while k < n:
            y, x = rng.uniform(border, shape[0] - border), rng.uniform(border, shape[1] - border)
            if k == 0 or np.min((ys[:k] - y) ** 2 + (xs[:k] - x) ** 2) >= min_sep * min_sep:
                ys[k], xs[k] = y, x
                k += 1
    else:
        ys = rng.uniform(border, shape[0] - border, n)
        xs = rng.uniform(border, shape[1] - border, n)
    lo, hi = flux
    u = rng.random(n)
    q = 1 - 2.0
    fs = (lo ** q + u * (hi ** q - lo ** q)) ** (1 / q)
    model = render_gauss(shape, ys, xs, fs, fwhm, radius) + bkg
    img = rng.poisson(model).astype(np.float64) + rng.normal(0, read_noise, shape)
    return img - bkg, ys, xs, fs
