"""Python wrappers around the oracle's detection functions (checker only)."""
import ctypes as C

import numpy as np

from _support import f64, load_pkg, oracle, ptr

abi = load_pkg()._abi
BORDER = {"replicate": abi.CP_BORDER_REPLICATE, "zero": abi.CP_BORDER_ZERO}


def _col(a):
    a = np.asarray(a, np.float64)
    return f64(a.T.ravel()), a.shape


def correlate(img, kernel, border="replicate"):
    a, (ny, nx) = _col(img)
    k, (kr, kc) = _col(kernel)
    out = np.zeros(ny * nx)
    rc = oracle().or_correlate(ptr(a, abi.f64p), ny, nx, ptr(k, abi.f64p), kr, kc, BORDER[border], ptr(out, abi.f64p))
    if rc:
        raise ValueError(oracle().or_last_error().decode())
    return out.reshape(nx, ny).T


def findlocalmaxima(img, edges=True):
    a, (ny, nx) = _col(img)
    cap = ny * nx + 1
    py = np.zeros(cap, np.int32)
    px = np.zeros(cap, np.int32)
    n = C.c_int32(0)
    rc = oracle().or_findlocalmaxima(ptr(a, abi.f64p), ny, nx, int(edges), cap, ptr(py, abi.i32p),
                                     ptr(px, abi.i32p), C.byref(n))
    if rc:
        raise ValueError(oracle().or_last_error().decode())
    return list(zip(py[: n.value].tolist(), px[: n.value].tolist()))


class MF:
    pass


def matched_filter(image, kernel, inv_var=None, normalize_zerosum=True, sigma=5.0, border="replicate"):
    a, (ny, nx) = _col(image)
    k, (kr, kc) = _col(kernel)
    iv = None if inv_var is None else _col(inv_var)[0]
    P = ny * nx
    cap = P + 1
    r = MF()
    r.smoothed_image = np.zeros(P)
    r.significance_map = np.zeros(P)
    r.smoothed_inv_var = np.zeros(P) if inv_var is not None else None
    r.kernel = np.zeros(kr * kc)
    py, px = np.zeros(cap, np.int32), np.zeros(cap, np.int32)
    ps, pf = np.zeros(cap), np.zeros(cap)
    s = abi.MFResult()
    s.smoothed = ptr(r.smoothed_image, abi.f64p)
    s.significance = ptr(r.significance_map, abi.f64p)
    s.smoothed_inv_var = ptr(r.smoothed_inv_var, abi.f64p)
    s.kernel = ptr(r.kernel, abi.f64p)
    s.peaks_cap = cap
    s.peak_y, s.peak_x = ptr(py, abi.i32p), ptr(px, abi.i32p)
    s.peak_significance, s.peak_flux = ptr(ps, abi.f64p), ptr(pf, abi.f64p)
    rc = oracle().or_matched_filter(ptr(a, abi.f64p), ptr(iv, abi.f64p), ny, nx, ptr(k, abi.f64p), kr, kc,
                                    int(normalize_zerosum), float(sigma), BORDER[border], C.byref(s))
    if rc:
        raise ValueError(oracle().or_last_error().decode())
    n = s.n_peaks
    shape = lambda v: None if v is None else v.reshape(nx, ny).T  # noqa: E731
    r.smoothed_image = shape(r.smoothed_image)
    r.significance_map = shape(r.significance_map)
    r.smoothed_inv_var = shape(r.smoothed_inv_var)
    r.kernel = r.kernel.reshape(kc, kr).T
    r.kernel_norm = s.kernel_norm
    r.peaks = list(zip(py[:n].tolist(), px[:n].tolist()))
    r.peak_significances = ps[:n].copy()
    r.peak_fluxes = pf[:n].copy()
    return r
