"""ctypes mirror of include/crowdphot_hip.h and the loader of libcrowdphot_hip.so.

The structures here must match the C header field for field.  The product
library is loaded from this package directory (``libcrowdphot_hip.so``, built
in-tree by ``build.py``); there is no fallback: a missing library raises.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_NAME = "libcrowdphot_hip.so"
LIB_PATH = os.path.join(HERE, LIB_NAME)

CP_OK = 0
CP_EARG = -1
CP_EHIP = -2
CP_ENOMEM = -3
CP_ESTATE = -4

CP_PSF_CIRCULAR_GAUSSIAN = 0
CP_PSF_CIRCULAR_MOFFAT = 1
CP_PSF_GRIDDED_IMAGE = 2

CP_SOLVER_LSQR = 0
CP_SOLVER_CGLS = 1
CP_SOLVER_LSMR = 2

CP_PREC_MIXED = 0
CP_BORDER_REPLICATE = 0
CP_BORDER_ZERO = 1
CP_PREC_FP64 = 1
CP_LAYOUT_COLMAJOR = 0
CP_LAYOUT_ROWMAJOR = 1

CP_COV_DEFAULT = 0
CP_COV_KNOWN = 1
CP_COV_REWEIGHTED = 2
CP_COV_NONE = -1

CP_FREE_Y = 1
CP_FREE_X = 2
CP_FREE_FLUX = 4

CP_DAMP_MARQUARDT = 0
CP_DAMP_LEVENBERG = 1
CP_DAMP_NONE = 2

NPARAMS = {CP_PSF_CIRCULAR_GAUSSIAN: 5, CP_PSF_CIRCULAR_MOFFAT: 6, CP_PSF_GRIDDED_IMAGE: 4}

f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)
i32p = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)
u8p = C.POINTER(C.c_uint8)


class PsfDesc(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("fwhm", C.c_double),
        ("alpha", C.c_double),
        ("beta", C.c_double),
        ("n_nodes", C.c_int32),
        ("gny", C.c_int32),
        ("gnx", C.c_int32),
        ("grids", f32p),
        ("os_x", C.c_int32),
        ("os_y", C.c_int32),
        ("origin_y", C.c_double),
        ("origin_x", C.c_double),
        ("fill_value", C.c_double),
        ("node_y", f64p),
        ("node_x", f64p),
    ]


class SimOpts(C.Structure):
    _fields_ = [
        ("solver", C.c_int32),
        ("inner_iterations", C.c_int32),
        ("linear_tol", C.c_double),
        ("model_rad", C.c_double),
        ("model_rad_max", C.c_double),
        ("model_rad_nsigma", C.c_double),
        ("max_step", C.c_double),
        ("max_trials", C.c_int32),
        ("max_iter", C.c_int32),
        ("x_tol", C.c_double),
        ("f_tol", C.c_double),
        ("g_tol", C.c_double),
        ("lambda_init", C.c_double),
        ("lambda_up", C.c_double),
        ("lambda_down", C.c_double),
        ("lambda_min", C.c_double),
        ("lambda_max", C.c_double),
        ("covariance", C.c_int32),
        ("free_mask", C.c_int32),
        ("fixed_y", C.c_double),
        ("fixed_x", C.c_double),
        ("fixed_flux", C.c_double),
        ("fixed_bkg", C.c_double),
        ("precision", C.c_int32),
        ("tile", C.c_int32),
        ("layout", C.c_int32),
        ("spread_model_fwhm", C.c_double),
    ]


class SimResult(C.Structure):
    _fields_ = [
        ("y", f64p), ("x", f64p), ("flux", f64p), ("bkg", f64p),
        ("y_err", f64p), ("x_err", f64p), ("flux_err", f64p), ("bkg_err", f64p),
        ("chisq", f64p),
        ("converged", u8p), ("valid", u8p),
        ("n_iter", i32p),
        ("active", u8p),
        ("residual", f32p),
        ("n_passes", C.c_int32), ("n_failed", C.c_int32),
        ("n_active", C.c_int32),
        ("n_union", C.c_int64),
        ("cost_final", C.c_double),
        ("n_trials_total", C.c_int32), ("n_inner_total", C.c_int32),
        ("trace", f64p),
        ("trace_cap", C.c_int32), ("trace_len", C.c_int32),
        ("qfit", f64p), ("qfit_expected", f64p), ("qfit_z", f64p), ("crowding", f64p),
        ("spread_model", f64p), ("spread_model_err", f64p),
    ]


class MFResult(C.Structure):
    _fields_ = [
        ("smoothed", f64p), ("smoothed_inv_var", f64p), ("significance", f64p), ("kernel", f64p),
        ("kernel_norm", C.c_double),
        ("peaks_cap", C.c_int32), ("n_peaks", C.c_int32),
        ("peak_y", i32p), ("peak_x", i32p),
        ("peak_significance", f64p), ("peak_flux", f64p),
    ]


class LMOpts(C.Structure):
    _fields_ = [
        ("lambda_init", C.c_double),
        ("lambda_up", C.c_double),
        ("lambda_down", C.c_double),
        ("lambda_min", C.c_double),
        ("lambda_max", C.c_double),
        ("max_iter", C.c_int32),
        ("x_tol", C.c_double),
        ("f_tol", C.c_double),
        ("g_tol", C.c_double),
        ("damping", C.c_int32),
        ("min_diagonal", C.c_double),
        ("covariance", C.c_int32),
    ]


class Tuning(C.Structure):
    """cp_tuning: blocks per compute unit of the grid-stride kernels and the
    simultaneous fit's work-tile sizes (render tile side, lanes per
    normal-matrix block, lanes per star of the fill) and the batched LM's work
    tile (cutouts per work-counter increment), and the largest Golub-Kahan grid
    whose step reductions fold into the next step; 0 = default."""
    _fields_ = [
        ("hbuild_blocks_per_cu", C.c_int32),
        ("cg_blocks_per_cu", C.c_int32),
        ("render_blocks_per_cu", C.c_int32),
        ("lm_blocks_per_cu", C.c_int32),
        ("render_tile", C.c_int32),
        ("hbuild_lanes", C.c_int32),
        ("fill_lanes", C.c_int32),
        ("lm_work_tile", C.c_int32),
        ("gk_fold_blocks", C.c_int32),
    ]


class LMStat(C.Structure):
    _fields_ = [
        ("minimum", C.c_double),
        ("cost_init", C.c_double),
        ("lambda_final", C.c_double),
        ("chisq", C.c_double),
        ("iterations", C.c_int32),
        ("converged", C.c_uint8),
        ("x_converged", C.c_uint8),
        ("f_converged", C.c_uint8),
        ("g_converged", C.c_uint8),
        ("status", C.c_int32),
    ]


def sim_opts_default():
    """Reference defaults of fit_all_stars_simultaneous
    (src/photometry/psf_photometry_simultaneous.jl:957-981)."""
    o = SimOpts()
    o.solver = CP_SOLVER_LSQR
    o.inner_iterations = 10
    o.linear_tol = 1e-6
    o.model_rad = 0.0
    o.model_rad_max = 0.0
    o.model_rad_nsigma = 1.0
    o.max_step = 0.25
    o.max_trials = 8
    o.max_iter = 40
    o.x_tol = 1e-5
    o.f_tol = 1e-4
    o.g_tol = 1e-4
    o.lambda_init = 1e-4
    o.lambda_up = 10.0
    o.lambda_down = 10.0
    o.lambda_min = 1e-12
    o.lambda_max = 1e12
    o.covariance = CP_COV_DEFAULT
    o.free_mask = CP_FREE_Y | CP_FREE_X | CP_FREE_FLUX
    o.fixed_y = o.fixed_x = o.fixed_flux = o.fixed_bkg = 0.0
    o.precision = CP_PREC_MIXED
    o.layout = CP_LAYOUT_COLMAJOR
    o.tile = 0
    return o


def lm_opts_default():
    """Reference defaults of lm_irls (src/levenberg_marquardt.jl:431-448)."""
    o = LMOpts()
    o.lambda_init = 1e-4
    o.lambda_up = 10.0
    o.lambda_down = 10.0
    o.lambda_min = 1e-12
    o.lambda_max = 1e12
    o.max_iter = 200
    o.x_tol = 1e-8
    o.f_tol = 1e-4
    o.g_tol = 1e-4
    o.damping = CP_DAMP_MARQUARDT
    o.min_diagonal = 1e-6
    o.covariance = CP_COV_DEFAULT
    return o


class Request(C.Structure):
    _fields_ = [
        ("psf", C.POINTER(PsfDesc)),
        ("image", C.c_void_p),
        ("inv_var", C.c_void_p),
        ("ny", C.c_int32), ("nx", C.c_int32),
        ("y0", C.c_void_p), ("x0", C.c_void_p), ("flux0", C.c_void_p),
        ("n_stars", C.c_int32),
        ("fit_rad", C.c_double),
        ("opts", C.POINTER(SimOpts)),
        ("out", C.POINTER(SimResult)),
        ("device_inputs", C.c_int32),
    ]


class RequestStat(C.Structure):
    _fields_ = [
        ("rc", C.c_int32),
        ("slot", C.c_int32),
        ("t_submit", C.c_double), ("t_start", C.c_double), ("t_end", C.c_double),
        ("hbm_bytes", C.c_uint64),
        ("hbm_charged", C.c_uint64),
        ("hbm_peak", C.c_uint64),
        ("hbm_resident_peak", C.c_uint64),
    ]


class MemStats(C.Structure):
    _fields_ = [("budget", C.c_uint64), ("in_use", C.c_uint64), ("cached", C.c_uint64), ("peak", C.c_uint64),
                ("peak_resident", C.c_uint64)]


# Every symbol include/crowdphot_hip.h declares, with (restype, argtypes).
SIGNATURES = {
    "cp_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.c_uint64]),
    "cp_destroy": (None, [C.c_void_p]),
    "cp_last_error": (C.c_char_p, [C.c_void_p]),
    "cp_stream": (C.c_void_p, [C.c_void_p]),
    "cp_synchronize": (C.c_int, [C.c_void_p]),
    "cp_profile": (C.c_int, [C.c_void_p, C.c_int]),
    "cp_get_tuning": (C.c_int, [C.c_void_p, C.POINTER(Tuning)]),
    "cp_set_tuning": (C.c_int, [C.c_void_p, C.POINTER(Tuning)]),
    "cp_profile_report": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int32]),
    "cp_mem_stats_get": (C.c_int, [C.c_void_p, C.POINTER(MemStats)]),
    "cp_set_budget": (C.c_int, [C.c_void_p, C.c_uint64]),
    "cp_release_cached": (C.c_int, [C.c_void_p]),
    "cp_sim_opts_default": (None, [C.POINTER(SimOpts)]),
    "cp_lm_opts_default": (None, [C.POINTER(LMOpts)]),
    "cp_psf_nparams": (C.c_int, [C.c_int32]),
    "cp_evaluate_fg": (C.c_int, [C.c_void_p, C.POINTER(PsfDesc), f64p, f64p, f64p, C.c_int64, f64p, f64p]),
    "cp_fit_simultaneous": (C.c_int, [C.c_void_p, C.POINTER(PsfDesc), f32p, f32p, C.c_int32, C.c_int32,
                                      f64p, f64p, f64p, C.c_int32, C.c_double, C.POINTER(SimOpts),
                                      C.POINTER(SimResult)]),
    "cp_fit_simultaneous_device": (C.c_int, [C.c_void_p, C.POINTER(PsfDesc), C.c_void_p, C.c_void_p,
                                             C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                             C.c_int32, C.c_double, C.POINTER(SimOpts),
                                             C.POINTER(SimResult)]),
    "cp_build_stamps": (C.c_int, [C.c_void_p, f32p, C.c_int32, C.c_int32, f64p, f64p, f64p, C.c_int32,
                                  C.c_double, i32p, i32p, i32p, i32p, i32p, u8p]),
    "cp_fit_all_stars_device": (C.c_int, [C.c_void_p, C.POINTER(PsfDesc), C.c_void_p, C.c_void_p, C.c_int32,
                                          C.c_int32, C.c_void_p, C.c_int32, C.c_double, C.c_int32, C.c_uint32,
                                          f64p, C.POINTER(LMOpts), C.c_double, C.c_int32, C.POINTER(SimResult)]),
    "cp_fit_all_stars": (C.c_int, [C.c_void_p, C.POINTER(PsfDesc), f32p, f32p, C.c_int32, C.c_int32, f64p,
                                   C.c_int32, C.c_double, C.c_int32, C.c_uint32, f64p, C.POINTER(LMOpts),
                                   C.c_double, C.c_int32, C.POINTER(SimResult)]),
    "cp_fit_star_batch": (C.c_int, [C.c_void_p, C.POINTER(PsfDesc), f32p, i64p, i32p, i32p, f32p, f64p,
                                    C.c_uint32, C.c_int32, C.POINTER(LMOpts), f64p, f64p,
                                    C.POINTER(LMStat)]),
    "cp_fit_star_batch_device": (C.c_int, [C.c_void_p, C.POINTER(PsfDesc), C.c_void_p, C.c_void_p,
                                           C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32,
                                           C.c_int32, C.POINTER(LMOpts), C.c_void_p, C.c_void_p,
                                           C.c_void_p]),
    "cp_fit_star_batch_device_noalloc": (C.c_int, [C.c_void_p, C.POINTER(PsfDesc), C.c_void_p, C.c_void_p,
                                                   C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32,
                                                   C.c_int32, C.POINTER(LMOpts), C.c_void_p, C.c_void_p,
                                                   C.c_void_p]),
    "cp_correlate": (C.c_int, [C.c_void_p, f64p, C.c_int32, C.c_int32, f64p, C.c_int32, C.c_int32, C.c_int32,
                               f64p]),
    "cp_correlate_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, f64p, C.c_int32, C.c_int32,
                                      C.c_int32, C.c_void_p]),
    "cp_findlocalmaxima": (C.c_int, [C.c_void_p, f64p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, i32p, i32p,
                                     C.POINTER(C.c_int32)]),
    "cp_matched_filter": (C.c_int, [C.c_void_p, f32p, f32p, C.c_int32, C.c_int32, f64p, C.c_int32, C.c_int32,
                                    C.c_int32, C.c_double, C.c_int32, C.POINTER(MFResult)]),
    "cp_matched_filter_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, f64p,
                                           C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_int32,
                                           C.POINTER(MFResult)]),
    "cp_sim_sources": (C.c_int, [C.c_uint64, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_double,
                                 C.c_int32, C.c_double, C.c_double, C.c_double, f64p, f64p, f64p]),
    "cp_sim_sources_plummer": (C.c_int, [C.c_uint64, C.c_int32, C.c_double, C.c_double, C.c_double,
                                         C.c_double, C.c_double, C.c_int32, C.c_double, C.c_double,
                                         C.c_double, f64p, f64p, f64p]),
    "cp_render_sources": (C.c_int, [C.POINTER(PsfDesc), C.c_int32, C.c_int32, f64p, f64p, f64p,
                                    C.c_int32, C.c_double, f64p]),
    "cp_add_noise": (C.c_int, [C.c_uint64, f64p, C.c_int64, C.c_double, C.c_double]),
    "cp_spread_kernel": (C.c_int, [C.c_double, C.c_int32, f64p]),
    "cp_curve_of_growth": (C.c_int, [C.POINTER(PsfDesc), C.c_int32, f64p, f64p]),
    "cp_pixel_overlap": (C.c_double, [C.c_double, C.c_double, C.c_double]),
    "cp_request_bytes": (C.c_uint64, [C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_int32,
                                      C.c_int32]),
    "cp_server_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int32, C.c_uint64, C.c_int32]),
    "cp_server_submit": (C.c_int, [C.c_void_p, C.POINTER(Request), C.POINTER(C.c_uint64)]),
    "cp_server_poll": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_int32)]),
    "cp_server_wait": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(RequestStat), C.c_char_p, C.c_int32]),
    "cp_server_last_error": (C.c_char_p, [C.c_void_p]),
    "cp_server_destroy": (None, [C.c_void_p]),
}

_lib = None


def load_library(path=None):
    """Load libcrowdphot_hip.so and bind every declared symbol.  Raises if absent."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"{p} is missing: build the HIP library first (python crowdphot.jl_amd/build.py)")
    lib = C.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib
