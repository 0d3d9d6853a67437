/*
 * oracle.h -- FP64 CPU restatement of CrowdPhot.jl's PSF-fitting hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or the timed CPU baseline).  The product (libcrowdphot_hip.so) never links,
 * loads or calls it.
 *
 * Every function restates the reference algorithm in plain float64 C++ and
 * cites the reference file:line it follows (paths relative to the
 * cgarling/CrowdPhot.jl tree).  Struct layouts come from the public header
 * include/crowdphot_hip.h so that the oracle and the product take identical
 * inputs.
 */
#ifndef CROWDPHOT_ORACLE_H
#define CROWDPHOT_ORACLE_H
#include "../include/crowdphot_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* models */
int or_nparams(int32_t kind);
int or_evaluate_fg(const cp_psf_desc* psf, const double* prm, double py, double px,
                   double* f, double* G);
double or_evaluate(const cp_psf_desc* psf, const double* prm, double py, double px);
int or_spread_kernel(double fwhm, int32_t half, double* K);
void or_bicubic(const double* data, int32_t ny, int32_t nx, double y, double x, double fill,
                double* out3);
int or_grid_corners_dw(const cp_psf_desc* psf, double Y, double X, int32_t* idx4, double* w4,
                       double* dwdy4, double* dwdx4);
double or_curve_of_growth(const cp_psf_desc* psf, double r);
double or_curve_of_growth_pixels(const cp_psf_desc* psf, double r);
double or_pixel_overlap(double dx, double dy, double r);

/* binning / indexing */
uint64_t or_morton2d(int64_t y, int64_t x);
int or_ring_offsets(int32_t R, int32_t* dy, int32_t* dx);
int or_model_radii(const cp_psf_desc* psf, double model_rad, double nsigma, int32_t R_fit,
                   int32_t R_cap, const double* w, int64_t npix, const double* flux_init,
                   int32_t n, int32_t* out);
int or_build_stamps(const double* w, int32_t ny, int32_t nx, const double* y0, const double* x0,
                    const double* flux0, int32_t n_stars, double fit_rad, int32_t* anchor_y,
                    int32_t* anchor_x, int32_t* active, int32_t* n_active, int32_t* pixels,
                    uint8_t* union_mask);

/* matrix-free operators on explicit stamp values V (p, S2, n_active) and
 * pixels (S2, n_active), 1-based flat indices, 0 = masked */
void or_apply_J(double* y, int64_t npix, const double* V, const int32_t* pixels, int32_t p,
                int32_t S2, int32_t n_active, const double* v, const uint8_t* live);
void or_apply_JT(double* z, const double* V, const int32_t* pixels, int32_t p, int32_t S2,
                 int32_t n_active, const double* u, const uint8_t* live);
/* damped least squares min ||J d - b||^2 + lam^2 ||d||^2 on the stamp operator.
 * solver: CP_SOLVER_LSQR or CP_SOLVER_CGLS.  Returns iterations run. */
int or_inner_solve(int32_t solver, const double* V, const int32_t* pixels, int32_t p, int32_t S2,
                   int32_t n_active, int64_t npix, const uint8_t* live, const double* b,
                   double lam, int32_t itmax, double atol, double btol, double* x_out);

/* whole fit (fit_all_stars_simultaneous) on a float64 image */
int or_fit_simultaneous(const cp_psf_desc* psf, const double* image, const double* inv_var,
                        int32_t ny, int32_t nx, const double* y0, const double* x0,
                        const double* flux0, int32_t n_stars, double fit_rad,
                        const cp_sim_opts* opts, cp_sim_result* out);
const char* or_last_error(void);

/* per-star LM */
int or_lm_line(const double* xs, const double* ys, int32_t n, const double* x0,
               const cp_lm_opts* o, double* x_out, cp_lm_stat* st);
int or_fit_star(const cp_psf_desc* psf, const double* pix, int32_t ny, int32_t nx,
                const double* inv_var, const double* theta0, uint32_t free_mask,
                const cp_lm_opts* o, double* theta_out, double* cov_out, cp_lm_stat* st);
int or_fit_all_stars(const cp_psf_desc* psf, const double* image, const double* inv_var, int32_t ny, int32_t nx,
                     const double* theta0, int32_t n, double fit_rad, int32_t n_passes, uint32_t free_mask,
                     const double* fixed_values, const cp_lm_opts* o, double spread_model_fwhm,
                     cp_sim_result* out);
/* batched over float32 cutouts packed as in cp_fit_star_batch; nthreads <= 0 = all */
int or_fit_star_batch(const cp_psf_desc* psf, const float* pix, const int64_t* offsets,
                      const int32_t* ny, const int32_t* nx, const float* inv_var,
                      const double* theta0, uint32_t free_mask, int32_t n_cutouts,
                      const cp_lm_opts* o, double* theta_out, double* cov_out, cp_lm_stat* st,
                      int32_t nthreads);
int or_extract_errors(const double* cov, int32_t p, const int32_t* free_idx, int32_t n_params,
                      double* errors_col);

/* detection */
int or_correlate(const double* img, int32_t ny, int32_t nx, const double* kernel, int32_t kr, int32_t kc,
                 int32_t border, double* out);
int or_findlocalmaxima(const double* img, int32_t ny, int32_t nx, int32_t edges, int32_t cap, int32_t* py,
                       int32_t* px, int32_t* n);
int or_matched_filter(const double* image, const double* inv_var, int32_t ny, int32_t nx, const double* kernel,
                      int32_t kr, int32_t kc, int32_t normalize_zerosum, double sigma, int32_t border,
                      cp_mf_result* out);

#ifdef __cplusplus
}
#endif
#endif
