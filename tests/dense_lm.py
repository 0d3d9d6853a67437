"""A dense restatement of fit_all_stars_simultaneous's outer loop for a few
CircularGaussianPSF stars (y, x, flux free), in numpy, used to pin the
oracle's trial / damping / freeze control flow (test infrastructure).

It follows src/photometry/psf_photometry_simultaneous.jl:1186-1284 (outer
loop, trials, accept, global g-cosine and f tests), :700-716 (step cap),
:739-823 (per-star freeze, the local x and g tests), :342-399 (equilibrated
stamp Jacobian) and :545-577 (model render), with the inner solve done
exactly: the damped least-squares step (J'J + lam I) d = -J'r by a dense
solve, which LSQR reaches when it is allowed enough iterations.

`variant` plants one deliberate mistake, so a test can show that the
comparison with the oracle would catch it:
  "lambda_up"  a rejected trial raises lambda by lambda_down instead of lambda_up
  "f_test"     the f test drops its cost_cand < cost condition
  "x_freeze"   the local x test measures the step and parameters without D
  "g_freeze"   the local g test uses the star's own reduced cost, not the global one
"""
import math

import numpy as np

GP = -4.0 * math.log(2.0)
EPS = np.finfo(float).eps


def _gauss(y, x, flux, fwhm, gy, gx):
    """f, df/dy, df/dx, df/dflux of CircularGaussianPSF (bkg 0) at pixels gy, gx."""
    f2 = fwhm * fwhm
    unit = -GP / (math.pi * f2)
    dy, dx = gy - y, gx - x
    g = np.exp(GP * (dx * dx + dy * dy) / f2)
    f = flux * unit * g
    return f, -2 * GP * f * dy / f2, -2 * GP * f * dx / f2, unit * g


def dense_fit(image, inv_var, y0, x0, f0, fwhm, fit_rad, model_rad, max_iter=40, max_trials=8, x_tol=1e-5,
              f_tol=1e-4, g_tol=1e-4, lambda_init=1e-4, lambda_up=10.0, lambda_down=10.0, lambda_min=1e-12,
              lambda_max=1e12, max_step=0.25, variant=None):
    ny, nx = image.shape
    n = len(y0)
    p = 3
    R = math.ceil(fit_rad)
    mR = min(max(math.ceil(model_rad), R), math.ceil(10 * fit_rad))
    ay, ax = np.rint(y0).astype(int), np.rint(x0).astype(int)
    w = np.asarray(inv_var, float)
    # stamps (dx-major), union pixels, model boxes
    stamps = []
    for j in range(n):
        pix = [(ay[j] + dy, ax[j] + dx) for dx in range(-R, R + 1) for dy in range(-R, R + 1)]
        stamps.append([(gy, gx) for gy, gx in pix if 1 <= gy <= ny and 1 <= gx <= nx and w[gy - 1, gx - 1] > 0])
    union = sorted(set(q for s in stamps for q in s), key=lambda q: (q[1], q[0]))
    uidx = {q: i for i, q in enumerate(union)}
    U = len(union)
    uy = np.array([q[0] for q in union], float)
    ux = np.array([q[1] for q in union], float)
    uw = np.array([w[q[0] - 1, q[1] - 1] for q in union])
    boxes = []
    for j in range(n):
        gy = np.arange(max(1, ay[j] - mR), min(ny, ay[j] + mR) + 1)
        gx = np.arange(max(1, ax[j] - mR), min(nx, ax[j] + mR) + 1)
        boxes.append(np.meshgrid(gy, gx, indexing="ij"))
    theta = np.column_stack([y0, x0, f0]).astype(float)
    data_work = np.asarray(image, float).copy()
    live = np.ones(n, bool)
    frozen_at = np.zeros(n, int)
    conv = np.zeros(n, bool)

    def model_on_union(th):
        m = np.zeros((ny, nx))
        for j in range(n):
            if live[j]:
                gy, gx = boxes[j]
                m[gy - 1, gx - 1] += _gauss(th[j, 0], th[j, 1], th[j, 2], fwhm, gy, gx)[0]
        return np.array([m[q[0] - 1, q[1] - 1] for q in union])

    def data_on_union():
        return np.array([data_work[q[0] - 1, q[1] - 1] for q in union])

    colnorm = np.zeros((n, p))
    Jhat = np.zeros((U, n * p))

    def fill():
        for j in range(n):
            if not live[j]:
                continue   # frozen columns keep their last values
            rows = np.array([uidx[q] for q in stamps[j]])
            gy, gx = uy[rows], ux[rows]
            _, dfy, dfx, dff = _gauss(theta[j, 0], theta[j, 1], theta[j, 2], fwhm, gy, gx)
            sw = np.sqrt(uw[rows])
            raw = np.column_stack([sw * dfy, sw * dfx, sw * dff])
            c = np.maximum(np.sqrt((raw ** 2).sum(0)), EPS)
            colnorm[j] = c
            Jhat[:, j * p:(j + 1) * p] = 0
            Jhat[rows, j * p:(j + 1) * p] = raw / c

    def jhat_live():
        J = Jhat.copy()
        for j in range(n):
            if not live[j]:
                J[:, j * p:(j + 1) * p] = 0
        return J

    fill()
    D = colnorm.copy()
    dof = max(U - n * p, 1)
    lam = lambda_init
    n_run = 0
    trace = []
    for it in range(1, max_iter + 1):
        fill()
        r = np.sqrt(uw) * (model_on_union(theta) - data_on_union())
        cost = float(np.sum(r * r))
        J = jhat_live()
        b_scaled = (J.T @ r).reshape(n, p)
        C_r = cost / dof
        gnorm = np.max(np.abs(colnorm * b_scaled) / (np.sqrt(colnorm ** 2 * C_r) + EPS))
        if gnorm <= g_tol:
            for j in range(n):
                if live[j]:
                    live[j], frozen_at[j], conv[j] = False, it - 1, True
            n_run = it - 1
            break
        accepted = f_conv = False
        for trial in range(1, max_trials + 1):
            d_scaled = -np.linalg.solve(J.T @ J + lam * np.eye(n * p), J.T @ r).reshape(n, p)
            d = d_scaled / colnorm
            ln = np.hypot(d[:, 0], d[:, 1])
            sc = np.where(ln > max_step, max_step / np.where(ln > 0, ln, 1), 1.0)
            d[:, 0] *= sc
            d[:, 1] *= sc
            d_scaled = colnorm * d
            rc = np.sqrt(uw) * (model_on_union(theta + d) - data_on_union())
            cost_c = float(np.sum(rc * rc))
            dcost = cost - cost_c
            if variant == "f_test":
                f_conv = dcost / cost <= f_tol
            else:
                f_conv = cost_c < cost and dcost / cost <= f_tol
            trace.append((it, trial, cost, cost_c, lam, int(live.sum())))
            if cost_c < cost:
                accepted = True
                theta = theta + d
                cost = cost_c
                lam = max(lam / lambda_down, lambda_min)
                D = np.maximum(D, colnorm)
                break
            lam = min(lam * (lambda_down if variant == "lambda_up" else lambda_up), lambda_max)
        n_run = it
        if f_conv:
            for j in range(n):
                if live[j]:
                    live[j], frozen_at[j], conv[j] = False, it, True
            break
        if accepted:
            for j in range(n):
                if not live[j]:
                    continue
                Dj = np.ones(p) if variant == "x_freeze" else D[j]
                v = (Dj / colnorm[j]) * d_scaled[j]
                xn = np.sqrt(np.sum((Dj * theta[j]) ** 2))
                thr = x_tol * (xn + x_tol)
                xt = np.sqrt(np.sum(v * v)) <= thr
                Cj = C_r
                if variant == "g_freeze":
                    rows = np.array([uidx[q] for q in stamps[j]])
                    Cj = float(np.sum(r[rows] ** 2)) / max(len(rows) - p, 1)
                gt = np.max(np.abs(colnorm[j] * b_scaled[j]) / (np.sqrt(colnorm[j] ** 2 * Cj) + EPS)) <= g_tol
                if xt or gt:
                    gy, gx = boxes[j]
                    data_work[gy - 1, gx - 1] -= _gauss(theta[j, 0], theta[j, 1], theta[j, 2], fwhm, gy, gx)[0]
                    live[j], frozen_at[j], conv[j] = False, it, True
            if not live.any():
                break
    n_iter = np.where(live, n_run, frozen_at)
    return dict(trace=np.array(trace, float), theta=theta, n_iter=n_iter, converged=conv, n_passes=n_run)
