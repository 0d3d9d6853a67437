// diagnostics.cpp -- host set-up of the spread_model reference disk for the
// post-fit diagnostics (src/photometry/psf_photometry_diagnostics.jl).  The
// per-star sums run on the device (k_diag).
#include "diag_host.h"

#include <cmath>

#include "host_model.h"

namespace cp {

double spread_fwhm_host(const cp_psf_desc* psf, double y0, double x0) {
    cp_host::HostModel M;
    if (!cp_host::make_model(psf, M) || !std::isfinite(y0) || !std::isfinite(x0)) return NAN;
    // extent(model) (PSF.jl:285-291 with fwhm_factor 5; ImagePSF empirical_models.jl:167-176;
    // every node of a GriddedPSFModel shares the grid geometry, gridded_psf.jl:281-298)
    double ylo, yhi, xlo, xhi;
    if (M.kind == CP_PSF_CIRCULAR_GAUSSIAN || M.kind == CP_PSF_CIRCULAR_MOFFAT) {
        const double fw = M.kind == CP_PSF_CIRCULAR_GAUSSIAN
                              ? M.fwhm
                              : 2 * M.alpha * std::sqrt(std::exp2(1 / M.beta) - 1);   // parametric_models.jl:689-692
        const double h = 5 * fw / 2;
        ylo = y0 - h;
        yhi = y0 + h;
        xlo = x0 - h;
        xhi = x0 + h;
    } else {
        ylo = y0 - (M.oy - 1) / M.osy;
        yhi = y0 + (M.gny - M.oy) / M.osy;
        xlo = x0 - (M.ox - 1) / M.osx;
        xhi = x0 + (M.gnx - M.ox) / M.osx;
    }
    const long iy0 = (long)std::floor(ylo), iy1 = (long)std::ceil(yhi);
    const long ix0 = (long)std::floor(xlo), ix1 = (long)std::ceil(xhi);
    if (iy1 < iy0 || ix1 < ix0 || (iy1 - iy0 + 1) * (ix1 - ix0 + 1) > (1L << 24)) return NAN;
    double s1 = 0, s2 = 0;
    for (long px = ix0; px <= ix1; ++px)
        for (long py = iy0; py <= iy1; ++py) {
            const double g = cp_host::eval_model(M, y0, x0, 1.0, (double)py, (double)px);
            s1 += g;
            s2 += g * g;
        }
    if (!(s2 > 0)) return NAN;
    const double ea = s1 * s1 / s2;
    return std::sqrt(ea * 2 * std::log(2.0) / M_PI);
}

// Mixture-of-Gaussians exponential disk (the reference's _EXPGAL_AMP / _EXPGAL_VAR,
// psf_photometry_diagnostics.jl:14-15).
static const double EXPGAL_AMP[6] = {1.99485977e-4, 2.61612679e-3, 1.89726655e-2,
                                     1.00186544e-1, 3.68534484e-1, 5.09490694e-1};
static const double EXPGAL_VAR[6] = {1.20078965e-3, 8.84526493e-3, 3.91463084e-2,
                                     1.39976817e-1, 4.60962500e-1, 1.50159566e0};

// I(m; v) = int_{-1/2}^{1/2} exp(-a k^2) cos(2 pi k m) dk, a = 2 pi^2 re^2 v:
// the band-limited 1-D profile of one mixture component, for m = 0..half at
// once.  The reference uses its closed form through the complex error
// function; here the (analytic, smooth) even integrand is integrated over
// [0, 1/2] by 16-point Gauss-Legendre on 32 equal panels (exact to float64
// rounding for |m| <= 31: each panel spans under 1/3 of a period of the
// fastest cosine), with cos(2 pi k m) from cos(2 pi k) by the Chebyshev
// recurrence.  The reference's underflow cut (pi^2 m^2 / a > 690 -> 0) is
// kept.
struct GaussLegendre16 {
    double x[16], w[16];
    GaussLegendre16() {   // nodes / weights of P_16 on [-1, 1] by Newton's method
        for (int i = 0; i < 16; ++i) {
            double z = std::cos(M_PI * (i + 0.75) / 16.5), dp = 1;
            for (int it = 0; it < 100; ++it) {
                double p0 = 1, p1 = 0;
                for (int k = 1; k <= 16; ++k) {
                    const double p2 = p1;
                    p1 = p0;
                    p0 = ((2 * k - 1) * z * p1 - (k - 1) * p2) / k;
                }
                dp = 16 * (z * p0 - p1) / (z * z - 1);
                const double dz = p0 / dp;
                z -= dz;
                if (std::fabs(dz) < 1e-16) break;
            }
            x[i] = z;
            w[i] = 2 / ((1 - z * z) * dp * dp);
        }
    }
};

static void band_profiles(double a, int half, double* Iv) {
    static const GaussLegendre16 gl;   // initialised once, thread-safe
    const double* X = gl.x;
    const double* Wt = gl.w;
    for (int m = 0; m <= half; ++m) Iv[m] = 0;
    const int panels = 32;
    const double hw = 0.25 / panels;   // panel half-width on [0, 1/2]
    for (int pnl = 0; pnl < panels; ++pnl) {
        const double mid = (2 * pnl + 1) * hw;
        for (int i = 0; i < 16; ++i) {
            const double k = mid + hw * X[i];
            const double e = Wt[i] * hw * std::exp(-a * k * k), c1 = std::cos(2 * M_PI * k);
            double cm1 = 1, cm = c1;   // cos(2 pi k (m - 1)), cos(2 pi k m)
            Iv[0] += e;
            for (int m = 1; m <= half; ++m) {
                Iv[m] += e * cm;
                const double nx = 2 * c1 * cm - cm1;
                cm1 = cm;
                cm = nx;
            }
        }
    }
    for (int m = 0; m <= half; ++m) Iv[m] = M_PI * M_PI * (double)m * m / a > 690 ? 0.0 : 2 * Iv[m];
}

void spread_kernel_host(double fwhm, int half, double* K) {
    const int n = 2 * half + 1;
    const double re = fwhm / 16 * 1.67834699;
    for (int i = 0; i < n * n; ++i) K[i] = 0;
    double Ih[32], Iv[64];
    for (int c = 0; c < 6; ++c) {
        const double a = 2 * M_PI * M_PI * re * re * EXPGAL_VAR[c];
        band_profiles(a, half, Ih);
        for (int m = -half; m <= half; ++m) Iv[m + half] = Ih[m < 0 ? -m : m];
        for (int jj = 0; jj < n; ++jj)
            for (int ii = 0; ii < n; ++ii) K[ii + jj * n] += EXPGAL_AMP[c] * Iv[ii] * Iv[jj];
    }
    double s = 0;
    for (int i = 0; i < n * n; ++i) s += K[i];
    for (int i = 0; i < n * n; ++i) K[i] /= s;
}

}  // namespace cp

// Test hook: the spread_model reference disk (for checks against the closed
// form); not part of the fitting ABI.
extern "C" int cp_spread_kernel(double fwhm, int32_t half, double* K) {
    if (!(fwhm > 0) || half < 0 || half > 31 || !K) return CP_EARG;
    cp::spread_kernel_host(fwhm, half, K);
    return CP_OK;
}
