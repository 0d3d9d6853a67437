"""Shared test helpers: load the product package and the FP64 oracle (checker)."""
import ctypes as C
import importlib.util
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "crowdphot.jl_amd")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def load_pkg():
    """Import the product package (its directory name is not an identifier)."""
    if "crowdphot_amd" in sys.modules:
        return sys.modules["crowdphot_amd"]
    spec = importlib.util.spec_from_file_location(
        "crowdphot_amd", os.path.join(PKG_DIR, "__init__.py"), submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["crowdphot_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def golden(name="reference_known_answers.json"):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


_oracle = None


def oracle():
    """ctypes handle on oracle/liboracle.so (FP64 restatement, checker only)."""
    global _oracle
    if _oracle is not None:
        return _oracle
    abi = load_pkg()._abi
    path = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(path):
        raise RuntimeError("oracle/liboracle.so missing: run `make -C oracle`")
    lib = C.CDLL(path)
    f64p, f32p, i32p, i64p, u8p = abi.f64p, abi.f32p, abi.i32p, abi.i64p, abi.u8p
    P = C.POINTER
    sig = {
        "or_nparams": (C.c_int, [C.c_int32]),
        "or_evaluate_fg": (C.c_int, [P(abi.PsfDesc), f64p, C.c_double, C.c_double, f64p, f64p]),
        "or_evaluate": (C.c_double, [P(abi.PsfDesc), f64p, C.c_double, C.c_double]),
        "or_spread_kernel": (C.c_int, [C.c_double, C.c_int32, f64p]),
        "or_bicubic": (None, [f64p, C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double, f64p]),
        "or_grid_corners_dw": (C.c_int, [P(abi.PsfDesc), C.c_double, C.c_double, i32p, f64p, f64p, f64p]),
        "or_curve_of_growth": (C.c_double, [P(abi.PsfDesc), C.c_double]),
        "or_curve_of_growth_pixels": (C.c_double, [P(abi.PsfDesc), C.c_double]),
        "or_pixel_overlap": (C.c_double, [C.c_double, C.c_double, C.c_double]),
        "or_morton2d": (C.c_uint64, [C.c_int64, C.c_int64]),
        "or_ring_offsets": (C.c_int, [C.c_int32, i32p, i32p]),
        "or_model_radii": (C.c_int, [P(abi.PsfDesc), C.c_double, C.c_double, C.c_int32, C.c_int32, f64p,
                                     C.c_int64, f64p, C.c_int32, i32p]),
        "or_build_stamps": (C.c_int, [f64p, C.c_int32, C.c_int32, f64p, f64p, f64p, C.c_int32, C.c_double,
                                      i32p, i32p, i32p, i32p, i32p, u8p]),
        "or_apply_J": (None, [f64p, C.c_int64, f64p, i32p, C.c_int32, C.c_int32, C.c_int32, f64p, u8p]),
        "or_apply_JT": (None, [f64p, f64p, i32p, C.c_int32, C.c_int32, C.c_int32, f64p, u8p]),
        "or_inner_solve": (C.c_int, [C.c_int32, f64p, i32p, C.c_int32, C.c_int32, C.c_int32, C.c_int64, u8p,
                                     f64p, C.c_double, C.c_int32, C.c_double, C.c_double, f64p]),
        "or_fit_simultaneous": (C.c_int, [P(abi.PsfDesc), f64p, f64p, C.c_int32, C.c_int32, f64p, f64p, f64p,
                                          C.c_int32, C.c_double, P(abi.SimOpts), P(abi.SimResult)]),
        "or_last_error": (C.c_char_p, []),
        "or_lm_line": (C.c_int, [f64p, f64p, C.c_int32, f64p, P(abi.LMOpts), f64p, P(abi.LMStat)]),
        "or_fit_star": (C.c_int, [P(abi.PsfDesc), f64p, C.c_int32, C.c_int32, f64p, f64p, C.c_uint32,
                                  P(abi.LMOpts), f64p, f64p, P(abi.LMStat)]),
        "or_fit_star_batch": (C.c_int, [P(abi.PsfDesc), f32p, i64p, i32p, i32p, f32p, f64p, C.c_uint32,
                                        C.c_int32, P(abi.LMOpts), f64p, f64p, P(abi.LMStat), C.c_int32]),
        "or_fit_all_stars": (C.c_int, [P(abi.PsfDesc), f64p, f64p, C.c_int32, C.c_int32, f64p, C.c_int32,
                                       C.c_double, C.c_int32, C.c_uint32, f64p, P(abi.LMOpts), C.c_double,
                                       P(abi.SimResult)]),
        "or_correlate": (C.c_int, [f64p, C.c_int32, C.c_int32, f64p, C.c_int32, C.c_int32, C.c_int32, f64p]),
        "or_findlocalmaxima": (C.c_int, [f64p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, i32p, i32p,
                                         C.POINTER(C.c_int32)]),
        "or_matched_filter": (C.c_int, [f64p, f64p, C.c_int32, C.c_int32, f64p, C.c_int32, C.c_int32, C.c_int32,
                                        C.c_double, C.c_int32, P(abi.MFResult)]),
        "or_extract_errors": (C.c_int, [f64p, C.c_int32, i32p, C.c_int32, f64p]),
    }
    for k, (r, a) in sig.items():
        fn = getattr(lib, k)
        fn.restype = r
        fn.argtypes = a
    _oracle = lib
    return lib


def ptr(a, t):
    return None if a is None else a.ctypes.data_as(t)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


_hip = None


def hip():
    """ctypes handle on the HIP runtime the product library itself links
    (libamdhip64.so.7), for tests that hand device buffers to the C ABI."""
    global _hip
    if _hip is None:
        load_pkg()._abi.load_library()   # resolves the same libamdhip64 the library uses
        h = C.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipDeviceSynchronize.argtypes = []
        _hip = h
    return _hip


class DeviceArray:
    """A device copy of a numpy array (hipMalloc + hipMemcpy); .ptr is the
    device address, .get() copies it back; freed on close / garbage collection."""

    def __init__(self, a=None, nbytes=None):
        h = hip()
        self.nbytes = int(a.nbytes if a is not None else nbytes)
        p = C.c_void_p()
        if h.hipMalloc(C.byref(p), max(self.nbytes, 1)) != 0:
            raise RuntimeError("hipMalloc failed")
        self.ptr = p
        if a is not None and self.nbytes:
            src = np.ascontiguousarray(a)
            if h.hipMemcpy(self.ptr, src.ctypes.data, self.nbytes, 1) != 0:   # host to device
                raise RuntimeError("hipMemcpy failed")

    def get(self, dtype, shape=None):
        out = np.empty(self.nbytes // np.dtype(dtype).itemsize, dtype)
        if self.nbytes and hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, 2) != 0:   # device to host
            raise RuntimeError("hipMemcpy failed")
        return out if shape is None else out.reshape(shape)

    def close(self):
        if getattr(self, "ptr", None):
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
