/*
 * crowdphot_hip.h -- C ABI of libcrowdphot_hip.so, the MI355X (gfx950) build of
 * CrowdPhot.jl's PSF-fitting hot path.
 *
 * Conventions (all entry points):
 *   - return 0 on success, a negative CP_E* code on failure; never throw.
 *     The message of the last failure is kept in cp_last_error(ctx).
 *   - images are column-major (y is the fastest index), 1-based coordinates:
 *     pixel (gy, gx) has flat index (gy-1) + (gx-1)*ny and its centre sits at
 *     (y, x) = (gy, gx), as in the reference (docs/src/index.md).
 *   - star parameters are float64; pixels, weights, PSF grids are float32.
 *   - a cp_ctx is not thread-safe; use one per host thread.  Every context
 *     owns one HIP stream (default priority) on its device.
 *
 * Reference interfaces each declaration replaces are cited as file:line of
 * cgarling/CrowdPhot.jl.
 */
#ifndef CROWDPHOT_HIP_H
#define CROWDPHOT_HIP_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ------------------------------------------------------- */
#define CP_OK 0
#define CP_EARG (-1)     /* ArgumentError in the reference */
#define CP_EHIP (-2)     /* HIP runtime failure */
#define CP_ENOMEM (-3)   /* HBM budget or allocation failure */
#define CP_ESTATE (-4)   /* bad handle / ticket */

/* ---- PSF models --------------------------------------------------------
 * CP_PSF_CIRCULAR_GAUSSIAN  src/psf/parametric_models.jl:62-113
 *     parameters (y, x, fwhm, flux, bkg)
 * CP_PSF_CIRCULAR_MOFFAT    src/psf/parametric_models.jl:672-724
 *     parameters (y, x, alpha, beta, flux, bkg)
 * CP_PSF_GRIDDED_IMAGE      GriddedPSFModel{T,<:ImagePSF{T}}
 *     src/psf/gridded_psf.jl:51-271 over ImagePSF src/psf/empirical_models.jl:76-318;
 *     a plain ImagePSF is the one-node case.  parameters (y, x, flux, bkg)
 */
#define CP_PSF_CIRCULAR_GAUSSIAN 0
#define CP_PSF_CIRCULAR_MOFFAT 1
#define CP_PSF_GRIDDED_IMAGE 2

/* Parameter slots in getproperties order, used by free masks. */
#define CP_PARAM_Y 0
#define CP_PARAM_X 1

typedef struct cp_psf_desc {
    int32_t kind;
    double fwhm;               /* CircularGaussianPSF.fwhm */
    double alpha, beta;        /* CircularMoffatPSF.alpha, .beta */
    /* ePSF nodes (kind == CP_PSF_GRIDDED_IMAGE); every node shares one geometry */
    int32_t n_nodes;
    int32_t gny, gnx;          /* rows (y) and columns (x) of each node grid, >= 4 */
    const float* grids;        /* n_nodes * gny * gnx, each node column-major */
    int32_t os_x, os_y;        /* ImagePSF.oversampling = (x factor, y factor) */
    double origin_y, origin_x; /* ImagePSF.origin (1-based grid coordinates) */
    double fill_value;         /* ImagePSF.fill_value */
    const double* node_y;      /* GriddedPSFModel.grid_y (n_nodes) */
    const double* node_x;      /* GriddedPSFModel.grid_x (n_nodes) */
} cp_psf_desc;

/* ---- simultaneous fit -----------------------------------------------------
 * Keyword arguments of fit_all_stars_simultaneous,
 * src/photometry/psf_photometry_simultaneous.jl:953-981, plus the build's
 * own solver / precision choices.  cp_sim_opts_default fills the reference
 * defaults. */
#define CP_SOLVER_LSQR 0       /* reference default (Krylov.jl :lsqr) */
#define CP_SOLVER_CGLS 1       /* CG on the damped normal equations (the build's own; J'J v fused) */
#define CP_SOLVER_LSMR 2       /* Krylov.jl :lsmr (Fong & Saunders 2011) */

#define CP_PREC_MIXED 0        /* FP32 pixels/Jacobian/residuals, FP64 reductions + star vectors */
#define CP_PREC_FP64 1         /* everything FP64 (parity runs) */

#define CP_LAYOUT_COLMAJOR 0   /* image[(gy-1) + (gx-1)*ny]: Julia's own layout */
#define CP_LAYOUT_ROWMAJOR 1   /* image[(gy-1)*nx + (gx-1)]: C / numpy order, transposed on the device */

#define CP_COV_DEFAULT 0       /* known weights if inv_var given, else reweighted */
#define CP_COV_KNOWN 1
#define CP_COV_REWEIGHTED 2
#define CP_COV_NONE (-1)

#define CP_FREE_Y 1
#define CP_FREE_X 2
#define CP_FREE_FLUX 4

typedef struct cp_sim_opts {
    int32_t solver;            /* CP_SOLVER_* */
    int32_t inner_iterations;  /* 10 */
    double linear_tol;         /* 1e-6, atol = btol of the inner solve */
    double model_rad;          /* > 0: scalar half-width; <= 0: :auto */
    double model_rad_max;      /* <= 0: 10 * fit_rad */
    double model_rad_nsigma;   /* 1.0 */
    double max_step;           /* 0.25 px */
    int32_t max_trials;        /* 8 */
    int32_t max_iter;          /* 40 */
    double x_tol, f_tol, g_tol;            /* 1e-5, 1e-4, 1e-4 */
    double lambda_init, lambda_up, lambda_down, lambda_min, lambda_max;
    int32_t covariance;        /* CP_COV_* */
    int32_t free_mask;         /* CP_FREE_* bits; fixed members take fixed_* values */
    double fixed_y, fixed_x, fixed_flux, fixed_bkg;  /* fixed values (bkg is always fixed) */
    int32_t precision;         /* CP_PREC_* */
    int32_t tile;              /* reserved, ignored: the render bins stars into 64 x 64 px tiles */
    int32_t layout;            /* CP_LAYOUT_*: image, inv_var and residual layout (default column-major) */
    double spread_model_fwhm;  /* > 0: FWHM of the spread_model reference disk; otherwise the
                                  Gaussian-equivalent FWHM of the PSF at the first catalogue star */
} cp_sim_opts;

/* Caller-allocated outputs, catalogue order (length n_stars unless noted).
 * Mirrors MultiPassPhotResult, src/photometry/psf_photometry_single.jl:99-122. */
typedef struct cp_sim_result {
    double *y, *x, *flux, *bkg;
    double *y_err, *x_err, *flux_err, *bkg_err;
    double *chisq;             /* may be NULL */
    uint8_t *converged, *valid;
    int32_t *n_iter;
    uint8_t *active;           /* may be NULL: 1 where the star entered the solve */
    float *residual;           /* ny*nx column-major, may be NULL */
    int32_t n_passes, n_failed;
    /* run statistics */
    int32_t n_active;
    int64_t n_union;           /* |union_pix| */
    double cost_final;
    int32_t n_trials_total, n_inner_total;
    double *trace;             /* optional: 8 doubles per trial (iter, trial, cost, cost_cand, lambda,
                                  n_live, inner iterations, the iteration's gradient cosine ||g||) */
    int32_t trace_cap, trace_len;
    /* post-fit diagnostics (psf_photometry_diagnostics.jl:154-231), each may be NULL;
       NaN where the reference leaves NaN */
    double *qfit, *qfit_expected, *qfit_z, *crowding, *spread_model, *spread_model_err;
} cp_sim_result;

/* ---- per-star Levenberg-Marquardt -----------------------------------------
 * lm_irls keyword arguments, src/levenberg_marquardt.jl:431-448 (no IRLS). */
#define CP_DAMP_MARQUARDT 0
#define CP_DAMP_LEVENBERG 1
#define CP_DAMP_NONE 2

typedef struct cp_lm_opts {
    double lambda_init, lambda_up, lambda_down, lambda_min, lambda_max;
    int32_t max_iter;          /* 200 */
    double x_tol, f_tol, g_tol;            /* 1e-8, 1e-4, 1e-4 */
    int32_t damping;           /* CP_DAMP_* */
    double min_diagonal;       /* MarquardtDamping.min_diagonal = 1e-6 */
    int32_t covariance;        /* CP_COV_* */
} cp_lm_opts;

/* Per-cutout outcome, the scalar fields of LMResult (src/levenberg_marquardt.jl:28-41). */
typedef struct cp_lm_stat {
    double minimum, cost_init, lambda_final, chisq;
    int32_t iterations;
    uint8_t converged, x_converged, f_converged, g_converged;
    int32_t status;            /* 0 ok, CP_EARG for a cutout the reference would reject */
} cp_lm_stat;

typedef struct cp_ctx cp_ctx;

/* ---- context -------------------------------------------------------------- */
/* New context on HIP device `device`; hbm_budget_bytes = 0 means no cap. */
int cp_create(cp_ctx** out, int device, uint64_t hbm_budget_bytes);
void cp_destroy(cp_ctx* ctx);
const char* cp_last_error(const cp_ctx* ctx);
/* The context's HIP stream (hipStream_t), for callers that order their own work on it. */
void* cp_stream(cp_ctx* ctx);
int cp_synchronize(cp_ctx* ctx);

/* Instrumentation, no reference counterpart: per-kernel HIP-event timing on
 * the context's stream.  cp_profile(ctx, 1) clears and starts it; the report
 * is one "name launches total_ms" line per kernel. */
int cp_profile(cp_ctx* ctx, int enable);
int cp_profile_report(cp_ctx* ctx, char* buf, int32_t cap);

/* Launch shapes, no reference counterpart: blocks per compute unit of the
 * grid-stride kernels (0 = the built-in default).  The autotuner
 * (crowdphot.jl_amd.autotune) picks them by timing whole calls.  They change
 * no result of the batched per-star LM or of the normal-matrix assembly; the
 * CG update's shape regroups the per-block partial sums of CG's norms
 * (rounding-level differences, fixed for a fixed shape).  Applies to calls made after cp_set_tuning. */
typedef struct cp_tuning {
    int32_t hbuild_blocks_per_cu;  /* normal-matrix assembly, 256-thread blocks (default 32) */
    int32_t cg_blocks_per_cu;      /* CG step update per star, 256-thread blocks (default 8) */
    int32_t render_blocks_per_cu;  /* tile-gather renders, 1024-thread blocks (default 2) */
    int32_t lm_blocks_per_cu;      /* batched per-star LM, 256-thread blocks (default 8) */
    /* work-tile sizes of the simultaneous fit */
    int32_t render_tile;           /* render tile side in pixels: 32 or 64 (default: 32 for images of fewer
                                      64-pixel tiles than CUs, else 64) */
    int32_t hbuild_lanes;          /* lanes per normal-matrix block: 4, 16 or 64 (default by stamp size) */
    int32_t fill_lanes;            /* lanes per star of the fill for stamps of <= 64 pixels: 16 or 64 (default 16) */
    /* work tile of the batched per-star LM */
    int32_t lm_work_tile;          /* cutouts a wavefront takes per work-counter increment: 1..64 (default 1) */
    /* Golub-Kahan steps of the simultaneous fit whose grid has at most this many blocks fold each step's
       reduction into the next step's launch (results identical either way): 1..1048576 (default: one
       block per CU) */
    int32_t gk_fold_blocks;
} cp_tuning;
int cp_get_tuning(cp_ctx* ctx, cp_tuning* out);
int cp_set_tuning(cp_ctx* ctx, const cp_tuning* t);

/* Device-memory accounting, no reference counterpart (Julia's GC owns the
 * reference's arrays): every buffer of a call comes from the context's
 * cache of device blocks.  in_use = blocks held by a running call, cached =
 * released blocks kept for reuse, peak = the most in_use and peak_resident =
 * the most in_use + cached since the last cp_set_budget.  cp_set_budget
 * changes the cap on in_use + cached (0 = none), frees cached blocks beyond
 * it and restarts both peaks; cp_release_cached frees every cached block.
 * The server uses them to keep each slot within its admission. */
typedef struct cp_mem_stats {
    uint64_t budget, in_use, cached, peak, peak_resident;
} cp_mem_stats;
int cp_mem_stats_get(cp_ctx* ctx, cp_mem_stats* out);
int cp_set_budget(cp_ctx* ctx, uint64_t hbm_budget_bytes);
int cp_release_cached(cp_ctx* ctx);

void cp_sim_opts_default(cp_sim_opts* o);
void cp_lm_opts_default(cp_lm_opts* o);

/* ---- PSF model evaluation (src/psf/PSF.jl:34-40,207-226) ------------------
 * Evaluates f and the full gradient G (getproperties order) of a model with
 * parameters prm (getproperties order) at n points (py[i], px[i]) on the
 * device.  G has nparams(kind) entries per point.  Host arrays. */
int cp_psf_nparams(int32_t kind);
int cp_evaluate_fg(cp_ctx* ctx, const cp_psf_desc* psf, const double* prm,
                   const double* py, const double* px, int64_t n,
                   double* f, double* G);

/* ---- simultaneous fit: fit_all_stars_simultaneous -------------------------
 * src/photometry/psf_photometry_simultaneous.jl:953-1431.
 * image / inv_var: ny*nx float32, column-major; inv_var may be NULL.
 * y0/x0/flux0: catalogue (n_stars).  Host pointers. */
int cp_fit_simultaneous(cp_ctx* ctx, const cp_psf_desc* psf,
                        const float* image, const float* inv_var, int32_t ny, int32_t nx,
                        const double* y0, const double* x0, const double* flux0, int32_t n_stars,
                        double fit_rad, const cp_sim_opts* opts, cp_sim_result* out);

/* Same, with image / inv_var / y0 / x0 / flux0 already resident in device
 * memory (device pointers); outputs in cp_sim_result are host pointers. */
int cp_fit_simultaneous_device(cp_ctx* ctx, const cp_psf_desc* psf,
                               const float* d_image, const float* d_inv_var, int32_t ny, int32_t nx,
                               const double* d_y0, const double* d_x0, const double* d_flux0,
                               int32_t n_stars, double fit_rad, const cp_sim_opts* opts,
                               cp_sim_result* out);

/* ---- star binning (integer path, exposed for bit-exact checks) ------------
 * _build_stamps! + _touched_pixels, psf_photometry_simultaneous.jl:245-335.
 * Host inputs (catalogue, weights ny*nx float32 or NULL = all ones).
 * Outputs (host, caller-allocated):
 *   anchor_y/anchor_x [n_stars] (round-half-even of y0/x0; INT32_MIN when not finite)
 *   active [n_stars]  catalogue ids (0-based) of active stars in Morton order, *n_active set
 *   pixels [S2 * n_stars] 1-based flat indices per active-order star, 0 = masked
 *   union_mask [ny*nx] 1 where the pixel is in union_pix
 * Stars whose anchor makes the reference's Morton key throw (negative
 * coordinates) or whose y/x are non-finite are excluded, as failures. */
int cp_build_stamps(cp_ctx* ctx, const float* w, int32_t ny, int32_t nx,
                    const double* y0, const double* x0, const double* flux0, int32_t n_stars,
                    double fit_rad, int32_t* anchor_y, int32_t* anchor_x,
                    int32_t* active, int32_t* n_active, int32_t* pixels, uint8_t* union_mask);

/* ---- batched per-star LM: fit_star over many cutouts ----------------------
 * fit_star src/psf/psf_fitting.jl:229-308 (+ the CircularGaussian / CircularMoffat /
 * ImagePSF / GriddedPSFModel specialisations :314-451, :772-862, :1086-1173,
 * :1177-1210) driving lm_irls src/levenberg_marquardt.jl:431-680.
 * Cutout c occupies pix[offsets[c] .. offsets[c] + ny[c]*nx[c]) column-major,
 * fit over all its pixels (inds = axes(cutout)), coordinates 1-based within
 * the cutout.  inv_var: NULL or packed like pix.  theta0 / theta_out:
 * n_cutouts * nparams(kind), getproperties order, fixed slots taken from
 * theta0.  cov_out: NULL or n_cutouts * p * p (p = popcount(free_mask)).
 * Host pointers; the _device variant takes pix / inv_var / offsets / ny / nx /
 * theta0 as device pointers and writes theta_out / cov_out / stats to device memory. */
int cp_fit_star_batch(cp_ctx* ctx, const cp_psf_desc* psf,
                      const float* pix, const int64_t* offsets, const int32_t* ny, const int32_t* nx,
                      const float* inv_var, const double* theta0, uint32_t free_mask,
                      int32_t n_cutouts, const cp_lm_opts* opts,
                      double* theta_out, double* cov_out, cp_lm_stat* stats);
int cp_fit_star_batch_device(cp_ctx* ctx, const cp_psf_desc* psf,
                             const float* d_pix, const int64_t* d_offsets,
                             const int32_t* d_ny, const int32_t* d_nx,
                             const float* d_inv_var, const double* d_theta0, uint32_t free_mask,
                             int32_t n_cutouts, const cp_lm_opts* opts,
                             double* d_theta_out, double* d_cov_out, cp_lm_stat* d_stats);

/* Allocation-free form of cp_fit_star_batch_device.  It has no counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, same results (the same kernels); the call
 * makes no device allocation: its work counter and the ePSF node grids (and,
 * for several nodes, their grid coordinates and index grid) go to one static
 * device array per GPU of CP_NOALLOC_SCRATCH_BYTES, whose address is looked up
 * once per device.  A PSF whose grids do not fit is rejected with CP_ENOMEM.
 * Every call on one GPU shares that array, so calls that overlap on one GPU
 * (other threads, other contexts' streams) race; serialise them.  Returns
 * when the results are in device memory. */
#define CP_NOALLOC_SCRATCH_BYTES (8u << 20)
int cp_fit_star_batch_device_noalloc(cp_ctx* ctx, const cp_psf_desc* psf,
                                     const float* d_pix, const int64_t* d_offsets,
                                     const int32_t* d_ny, const int32_t* d_nx,
                                     const float* d_inv_var, const double* d_theta0, uint32_t free_mask,
                                     int32_t n_cutouts, const cp_lm_opts* opts,
                                     double* d_theta_out, double* d_cov_out, cp_lm_stat* d_stats);

/* ---- sequential multi-pass photometry: fit_all_stars ---------------------
 * fit_all_stars, src/photometry/psf_photometry_single.jl:267-469 (replaces
 * that Julia function; every model kind above).
 * image / inv_var (NULL = unweighted): host float32 ny*nx column-major.
 * theta0: n_stars x nparams(kind) initial parameters in getproperties order
 * (_extract_source_catalog, :152-197: the catalogue's y, x, flux, bkg with
 * the template PSF's shape parameters).  free_mask: free parameters (bit k =
 * parameter k); a fixed parameter k fits at fixed_values[k] (the `fixed`
 * NamedTuple), or at the star's own value where fixed_values is NULL or NaN.
 * spread_model_fwhm <= 0: from the PSF at star 1 (_spread_fwhm).  precision:
 * CP_PREC_MIXED (float32 model / gradient, float64 sums and residual image)
 * or CP_PREC_FP64.  Fills y, x, flux, bkg, their errors (NaN for a free
 * parameter never fitted, 0 for a fixed one), chisq, converged, valid, n_iter
 * (summed over passes), the six diagnostics (NaN where not computed),
 * n_passes (0 for an empty catalogue), n_failed and, when set, residual.
 * Stars whose fit throws in the reference (too few valid inv_var pixels) are
 * marked invalid and counted in n_failed; a valid star with a non-finite
 * position fails the call with CP_EARG (the reference's InexactError). */
int cp_fit_all_stars(cp_ctx* ctx, const cp_psf_desc* psf, const float* image, const float* inv_var,
                     int32_t ny, int32_t nx, const double* theta0, int32_t n_stars, double fit_rad,
                     int32_t n_passes, uint32_t free_mask, const double* fixed_values,
                     const cp_lm_opts* opts, double spread_model_fwhm, int32_t precision,
                     cp_sim_result* out);

/* Same, with image / inv_var / theta0 already resident in device memory
 * (device pointers); outputs in cp_sim_result are host pointers. */
int cp_fit_all_stars_device(cp_ctx* ctx, const cp_psf_desc* psf, const float* d_image,
                            const float* d_inv_var, int32_t ny, int32_t nx, const double* d_theta0,
                            int32_t n_stars, double fit_rad, int32_t n_passes, uint32_t free_mask,
                            const double* fixed_values, const cp_lm_opts* opts,
                            double spread_model_fwhm, int32_t precision, cp_sim_result* out);

/* ---- detection: correlate, findlocalmaxima, matched_filter ----------------
 * correlate / correlate! (src/correlation.jl:108-140), findlocalmaxima and
 * matched_filter (src/detection.jl:87-324): the catalogue-building step before
 * the fits (_extract_source_catalog takes a MatchedFilterResult,
 * src/photometry/psf_photometry_single.jl:190-197).  Images are column-major
 * (row = y fastest); peak coordinates are 1-based (row, column). */
#define CP_BORDER_REPLICATE 0
#define CP_BORDER_ZERO 1

typedef struct cp_mf_result {
    double *smoothed;          /* ny*nx: image correlated with the normalised kernel; may be NULL */
    double *smoothed_inv_var;  /* ny*nx: inv_var correlated with the squared kernel (inv_var given); may be NULL */
    double *significance;      /* ny*nx significance map; may be NULL */
    double *kernel;            /* kr*kc normalised kernel; may be NULL */
    double kernel_norm;        /* out */
    int32_t peaks_cap;         /* capacity of the four peak arrays */
    int32_t n_peaks;           /* out: peaks found; only min(n_peaks, peaks_cap) are written */
    int32_t *peak_y, *peak_x;  /* 1-based row (y) and column (x) */
    double *peak_significance, *peak_flux;
} cp_mf_result;

/* out = img correlated with the kr x kc kernel (odd sides, column-major;
 * no flip), border CP_BORDER_REPLICATE or CP_BORDER_ZERO.  A kernel that is
 * numerically rank 1 (every singular value past the first below
 * sqrt(eps)) runs as two 1-D passes, as the reference's _tryfactor does.
 * Host float64 arrays; the _device variant takes device img / out. */
int cp_correlate(cp_ctx* ctx, const double* img, int32_t ny, int32_t nx, const double* kernel,
                 int32_t kr, int32_t kc, int32_t border, double* out);
int cp_correlate_device(cp_ctx* ctx, const double* d_img, int32_t ny, int32_t nx,
                        const double* kernel, int32_t kr, int32_t kc, int32_t border, double* d_out);
/* Pixels strictly greater than their 8 neighbours (3x3 window), in the
 * reference's order; edges = 0 skips the image boundary.  Host img; peak
 * arrays of capacity cap; *n_peaks = peaks found. */
int cp_findlocalmaxima(cp_ctx* ctx, const double* img, int32_t ny, int32_t nx, int32_t edges,
                       int32_t cap, int32_t* peak_y, int32_t* peak_x, int32_t* n_peaks);
/* matched_filter(image, kernel; inv_var, normalize_zerosum, sigma, border).
 * image / inv_var (NULL = uniform weights): float32 ny*nx; kernel: host
 * float64 kr*kc.  The _device variant takes device image / inv_var. */
int cp_matched_filter(cp_ctx* ctx, const float* image, const float* inv_var, int32_t ny, int32_t nx,
                      const double* kernel, int32_t kr, int32_t kc, int32_t normalize_zerosum,
                      double sigma, int32_t border, cp_mf_result* out);
int cp_matched_filter_device(cp_ctx* ctx, const float* d_image, const float* d_inv_var, int32_t ny,
                             int32_t nx, const double* kernel, int32_t kr, int32_t kc,
                             int32_t normalize_zerosum, double sigma, int32_t border, cp_mf_result* out);

/* ---- synthetic inputs: simulate_image (src/simulation.jl:211-498) ----------
 * Seeded xoshiro256** streams; host arrays.
 * cp_sim_sources: n positions uniform in [lo, hi] per axis and fluxes from
 *   flux_dist 0 = uniform, 1 = loguniform, 2 = powerlaw dN/dF ~ F^-flux_power
 *   (_sample_flux, src/simulation.jl:1-28).
 * cp_sim_sources_plummer: Plummer surface density of scale radius a about
 *   (cy, cx), rejecting points outside [lo, hi] per axis.
 * cp_render_sources: image[ny*nx] += each source over floor(y-r):ceil(y+r)
 *   (render_sources!, src/simulation.jl:321-347), float64 image.
 * cp_add_noise: Poisson(gain*v)/gain (Knuth below 30, rounded normal
 *   approximation above) then + N(0, read_noise^2) (add_noise!, :409-435). */
int cp_sim_sources(uint64_t seed, int32_t n, double ylo, double yhi, double xlo, double xhi,
                   int32_t flux_dist, double flo, double fhi, double flux_power,
                   double* y, double* x, double* flux);
int cp_sim_sources_plummer(uint64_t seed, int32_t n, double a, double cy, double cx,
                           double lo, double hi, int32_t flux_dist, double flo, double fhi,
                           double flux_power, double* y, double* x, double* flux);
int cp_render_sources(const cp_psf_desc* psf, int32_t ny, int32_t nx, const double* y,
                      const double* x, const double* flux, int32_t n, double model_radius,
                      double* image);
int cp_add_noise(uint64_t seed, double* image, int64_t npix, double gain, double read_noise);

/* The spread_model reference disk (_exp_disk_kernel_bandlimited,
 * psf_photometry_diagnostics.jl:78-101): K[i + j (2 half + 1)], sum 1.
 * Exposed so the diagnostics set-up can be checked on its own. */
int cp_spread_kernel(double fwhm, int32_t half, double* K);

/* curve_of_growth(model, radii) (src/curve_of_growth.jl:143-176) of the
 * unit-flux model centred at (0, 0) with exact pixel-circle overlap weights
 * (src/apertures.jl:52-166): what model_rad = :auto uses for ImagePSF and
 * GriddedPSFModel (_model_radii, psf_photometry_simultaneous.jl:209-243).
 * radii strictly positive and increasing.  Exposed for checks. */
int cp_curve_of_growth(const cp_psf_desc* psf, int32_t n_radii, const double* radii, double* flux);
/* Exact overlap area of the unit pixel centred at (dx, dy) with the circle of
 * radius r about the origin (aperture_weight(..., ExactOverlap()),
 * src/apertures.jl:160-166).  Exposed for checks. */
double cp_pixel_overlap(double dx, double dy, double r);

/* ---- serving: continuous batching of simultaneous fits ---------------------
 * No counterpart in the reference, which fits one image per call: for callers
 * that stream many image tiles.  A server owns n_slots contexts (one HIP
 * stream and device-block cache each) on one device.  Requests are admitted
 * in arrival order as soon as a slot is free and their HBM estimate
 * (cp_request_bytes) fits beside the running ones within hbm_budget (0 = no
 * cap); admitted fits overlap on the device.  Each result is the one the
 * request gets alone (its own buffers and stream; fixed-order reductions).
 * The caller keeps every array of a request (inputs, the psf grids and node
 * coordinates, the result arrays) alive until cp_server_wait returns for it.
 * HBM: a running request is charged its estimate (cp_request_bytes, an upper
 * bound on what the fit allocates) plus the blocks its slot kept cached from
 * earlier requests (which the fit reuses), and its slot context is capped
 * there; an idle slot is charged for its cached blocks.  Admission keeps the
 * sum over slots within hbm_budget, releasing idle slots' caches (then the
 * admitted slot's own) only when the next request needs the room.  cp_server_destroy finishes running
 * requests, completes queued ones with CP_ESTATE and returns once every
 * thread blocked in cp_server_wait has been woken; at most one thread may
 * wait on a ticket. */
typedef struct cp_server cp_server;

typedef struct cp_request {
    const cp_psf_desc* psf;     /* copied at submit */
    const float* image;         /* ny*nx column-major */
    const float* inv_var;       /* may be NULL */
    int32_t ny, nx;
    const double *y0, *x0, *flux0;
    int32_t n_stars;
    double fit_rad;
    const cp_sim_opts* opts;    /* copied at submit */
    cp_sim_result* out;         /* filled by the time cp_server_wait returns */
    int32_t device_inputs;      /* 1: image, inv_var, y0, x0, flux0 are device pointers */
} cp_request;

typedef struct cp_request_stat {
    int32_t rc;                 /* the fit's return code */
    int32_t slot;               /* worker slot that ran it */
    double t_submit, t_start, t_end;   /* seconds, one monotonic host clock */
    uint64_t hbm_bytes;         /* admission estimate (cp_request_bytes) */
    uint64_t hbm_charged;       /* bytes its slot held against the budget: the estimate plus the
                                   blocks the slot kept cached from earlier requests */
    uint64_t hbm_peak;          /* device blocks the fit held at most */
    uint64_t hbm_resident_peak; /* its slot's blocks (held + cached) at most, <= hbm_charged */
} cp_request_stat;

uint64_t cp_request_bytes(int32_t ny, int32_t nx, int32_t n_stars, double fit_rad, double model_rad,
                          int32_t precision, int32_t device_inputs);
int cp_server_create(cp_server** out, int32_t device, uint64_t hbm_budget, int32_t n_slots);
int cp_server_submit(cp_server* s, const cp_request* r, uint64_t* ticket);
int cp_server_poll(cp_server* s, uint64_t ticket, int32_t* done);
/* Blocks until the request finished; returns its rc (message in msg). */
int cp_server_wait(cp_server* s, uint64_t ticket, cp_request_stat* st, char* msg, int32_t msg_cap);
const char* cp_server_last_error(const cp_server* s);
void cp_server_destroy(cp_server* s);

#ifdef __cplusplus
}
#endif
#endif /* CROWDPHOT_HIP_H */
