// aperture.cpp -- exact pixel / circle overlap and the pixel-integrated curve
// of growth (src/apertures.jl, src/curve_of_growth.jl:143-176) used to size
// per-star model stamps with model_rad = :auto for ImagePSF / GriddedPSFModel
// (_model_radii, src/photometry/psf_photometry_simultaneous.jl:209-243).
#include "aperture_host.h"

#include <algorithm>
#include <cmath>
#include <vector>

#include "host_model.h"

namespace cp {
namespace {

// int_0^x sqrt(r^2 - t^2) dt, x clamped to [-r, r].
double chord_integral(double x, double r) {
    x = std::min(std::max(x, -r), r);
    // (r - x)(r + x): exactly 0 at x = +-r (r^2 - x^2 may be contracted to an
    // fma that leaves a rounding residue there)
    return 0.5 * (x * std::sqrt(std::max(0.0, (r - x) * (r + x))) + r * r * std::asin(x / r));
}

// Area of {(x, y): x >= a, y >= b, x^2 + y^2 <= r^2}.  With h(x) =
// sqrt(r^2 - x^2) the vertical extent at x is [max(b, -h), h]: for b >= 0 it
// is h - b on |x| < sqrt(r^2 - b^2); for b < 0 it is h - b on that interval
// and the full chord 2h outside it.
double quadrant_area(double a, double b, double r) {
    if (a >= r || b >= r) return 0.0;
    const double lo = std::max(a, -r);
    auto H = [&](double x) { return chord_integral(x, r); };
    if (b >= 0) {
        const double s = std::sqrt((r - b) * (r + b)), l = std::max(lo, -s);
        return l < s ? H(s) - H(l) - b * (s - l) : 0.0;
    }
    if (b <= -r) return 2 * (H(r) - H(lo));
    const double s = std::sqrt((r - b) * (r + b));
    double area = 0;
    auto full = [&](double x0, double x1) {
        x0 = std::max(x0, lo);
        if (x1 > x0) area += 2 * (H(x1) - H(x0));
    };
    full(-r, -s);
    {
        const double x0 = std::max(-s, lo);
        if (s > x0) area += H(s) - H(x0) - b * (s - x0);
    }
    full(s, r);
    return area;
}

// Area of the rectangle [x0, x1] x [y0, y1] inside the circle, by
// inclusion-exclusion over the four quadrant regions anchored at its corners.
double rect_area(double x0, double y0, double x1, double y1, double r) {
    return quadrant_area(x0, y0, r) - quadrant_area(x1, y0, r) - quadrant_area(x0, y1, r) +
           quadrant_area(x1, y1, r);
}

}  // namespace

double pixel_overlap(double dx, double dy, double r) {
    const double d2 = dx * dx + dy * dy, dr = std::sqrt(2.0) / 2;   // centre-to-corner distance
    if (d2 > (r + dr) * (r + dr)) return 0.0;
    if (r > dr && d2 < (r - dr) * (r - dr)) return 1.0;
    const double a = rect_area(dx - 0.5, dy - 0.5, dx + 0.5, dy + 0.5, r);
    return std::min(std::max(a, 0.0), 1.0);
}

bool model_cog_host(const cp_psf_desc* psf, int n_radii, const double* radii, double* flux) {
    cp_host::HostModel M;
    if (n_radii < 1 || !radii || !flux || !cp_host::make_model(psf, M)) return false;
    for (int k = 0; k < n_radii; ++k)
        if (!(radii[k] > 0) || !std::isfinite(radii[k]) || (k > 0 && !(radii[k] > radii[k - 1]))) return false;
    // bounding pixels of the aperture about (0, 0) (apertures.jl:52-65): a
    // lower edge that falls exactly on a pixel border starts one pixel in
    auto lo_edge = [](double v) {
        const double c = std::ceil(v);
        return (long)(c == v ? c + 1 : c);
    };
    auto hi_edge = [](double v) { return (long)std::ceil(v); };
    const double rmax = radii[n_radii - 1];
    const long L = std::max(-lo_edge(-rmax - 0.5), hi_edge(rmax - 0.5));
    if (L > 4096) return false;
    const long W = 2 * L + 1;
    // the model once per pixel of the largest box, reused for every radius
    std::vector<double> g((size_t)W * W);
    for (long j = -L; j <= L; ++j)
        for (long i = -L; i <= L; ++i)
            g[(size_t)(i + L) + (size_t)(j + L) * W] = cp_host::eval_model(M, 0.0, 0.0, 1.0, (double)i, (double)j);
    for (int k = 0; k < n_radii; ++k) {
        const double r = radii[k];
        const long i0 = lo_edge(-r - 0.5), i1 = hi_edge(r - 0.5);
        double s = 0;
        for (long j = i0; j <= i1; ++j)
            for (long i = i0; i <= i1; ++i) {
                const double w = pixel_overlap((double)j, (double)i, r);
                if (w > 0) s += w * g[(size_t)(i + L) + (size_t)(j + L) * W];
            }
        flux[k] = s;
    }
    return true;
}

}  // namespace cp

// Test hook: the curve of growth model_rad = :auto uses for ImagePSF /
// GriddedPSFModel; not part of the fitting ABI.
extern "C" int cp_curve_of_growth(const cp_psf_desc* psf, int32_t n_radii, const double* radii, double* flux) {
    return cp::model_cog_host(psf, n_radii, radii, flux) ? CP_OK : CP_EARG;
}

// Test hook: exact overlap of the unit pixel centred at (dx, dy) with the
// circle of radius r about the origin.
extern "C" double cp_pixel_overlap(double dx, double dy, double r) { return cp::pixel_overlap(dx, dy, r); }
