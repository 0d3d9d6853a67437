// oracle.cpp -- FP64 CPU restatement of CrowdPhot.jl's PSF-fitting hot path.
//
// TEST INFRASTRUCTURE ONLY (the checker and the timed CPU baseline); the
// product library never links or calls it.  Each function cites the
// reference file:line it restates (paths inside cgarling/CrowdPhot.jl).
//
// Parity status of each part (how it is pinned is listed in the tests):
//   models, bicubic, gridded corners ........ pinned (reference known answers,
//                                             closed forms, finite differences)
//   binning / Morton / ring / model radii .... pinned (closed-form definitions,
//                                             reference invariants)
//   J, J^T ................................... pinned (dense-J identities)
//   LSQR / LSMR / CG ......................... pinned to the damped normal
//       equations, to the exact Krylov-subspace minimisers at every k (LSQR
//       and CG: the residual norm; LSMR: the normal-equation residual) and,
//       for LSQR and LSMR, to scipy's implementations iterate by iterate;
//       Krylov.jl 0.10's own stopping-rule trajectory is parity unpinned
//       (the package is not available here).
//   outer Gauss-Newton loop .................. pinned by noiseless recovery and
//       frozen-star invariants; crowded-field trajectories follow the same
//       restated inner solver (Krylov.jl trajectory unpinned, as above).
//   lm_irls .................................. pinned (reference line-fit test,
//                                             noiseless recovery)
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <limits>
#include <numeric>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

thread_local std::string g_err;

int fail(const std::string& m) {
    g_err = m;
    return CP_EARG;
}

const double GAUSS_PRE = -4.0 * std::log(2.0);  // parametric_models.jl:18
const double EPS64 = std::numeric_limits<double>::epsilon();

// ---------------------------------------------------------------------------
// ePSF helpers
// ---------------------------------------------------------------------------

// Keys a=-1/2 cubic convolution on four samples, value and derivative at t
// (empirical_models.jl:185-196).
inline void cubic4(double f1, double f2, double f3, double f4, double t, double& val,
                   double& der) {
    double c1 = (f3 - f1) / 2;
    double c4 = f3 - f2 - c1;
    double c2 = 3 * c4 - (f4 - f2) / 2 + c1;
    double c3 = c4 - c2;
    double c4t = t * c3;
    val = t * (t * (c4t + c2) + c1) + f2;
    der = t * (3 * c4t + 2 * c2) + c1;
}

// bicubic_interpolate (empirical_models.jl:239-292): (value, d/dy, d/dx) at
// 1-based array coordinates (y = row, x = column) of a column-major ny x nx
// grid; outside the closed bounds -> (fill, 0, 0); stencil clamped to [1, n].
void bicubic(const double* data, int ny, int nx, double y, double x, double fill, double out[3]) {
    bool inb = std::isfinite(x) && std::isfinite(y) && x >= 1.0 && x <= nx && y >= 1.0 && y <= ny;
    double xc = inb ? x : 1.0, yc = inb ? y : 1.0;
    int lx = std::min(std::max((int)std::floor(xc), 1), nx - 1);
    int ly = std::min(std::max((int)std::floor(yc), 1), ny - 1);
    double dx = xc - lx, dy = yc - ly;
    int ix[4] = {std::max(lx - 1, 1), lx, lx + 1, std::min(lx + 2, nx)};
    int iy[4] = {std::max(ly - 1, 1), ly, ly + 1, std::min(ly + 2, ny)};
    double row[4], drow[4];
    for (int r = 0; r < 4; ++r) {
        const double* col = data + (iy[r] - 1);
        double f[4];
        for (int c = 0; c < 4; ++c) f[c] = col[(size_t)(ix[c] - 1) * ny];
        cubic4(f[0], f[1], f[2], f[3], dx, row[r], drow[r]);
    }
    double val, dfdy, dfdx, unused;
    cubic4(row[0], row[1], row[2], row[3], dy, val, dfdy);
    cubic4(drow[0], drow[1], drow[2], drow[3], dy, dfdx, unused);
    out[0] = inb ? val : fill;
    out[1] = inb ? dfdy : 0.0;
    out[2] = inb ? dfdx : 0.0;
}

struct Psf {
    int kind = 0;
    double fwhm = 0, alpha = 0, beta = 0;
    int n_nodes = 0, gny = 0, gnx = 0, os_x = 1, os_y = 1;
    double oy = 0, ox = 0, fill = 0;
    std::vector<double> data;            // n_nodes * gny * gnx
    std::vector<double> ygrid, xgrid;    // sorted unique node coordinates
    std::vector<int> index_grid;         // ny_g x nx_g column-major, 1-based node ids
};

// GriddedPSFModel constructor bookkeeping (gridded_psf.jl:68-144).
int make_psf(const cp_psf_desc* d, Psf& P) {
    if (!d) return fail("psf descriptor is NULL");
    P.kind = d->kind;
    if (d->kind == CP_PSF_CIRCULAR_GAUSSIAN) {
        P.fwhm = d->fwhm;
        return 0;
    }
    if (d->kind == CP_PSF_CIRCULAR_MOFFAT) {
        P.alpha = d->alpha;
        P.beta = d->beta;
        return 0;
    }
    if (d->kind != CP_PSF_GRIDDED_IMAGE) return fail("unknown psf kind");
    if (d->n_nodes < 1) return fail("`psfs` must be non-empty");
    if (d->gny < 4 || d->gnx < 4) return fail("each axis of `data` must have at least 4 samples");
    P.n_nodes = d->n_nodes;
    P.gny = d->gny;
    P.gnx = d->gnx;
    P.os_x = d->os_x;
    P.os_y = d->os_y;
    P.oy = d->origin_y;
    P.ox = d->origin_x;
    P.fill = d->fill_value;
    size_t nn = (size_t)d->gny * d->gnx;
    P.data.resize(nn * d->n_nodes);
    for (size_t i = 0; i < P.data.size(); ++i) P.data[i] = (double)d->grids[i];
    if (d->n_nodes == 1) {
        P.ygrid = {d->node_y ? d->node_y[0] : 0.0};
        P.xgrid = {d->node_x ? d->node_x[0] : 0.0};
        P.index_grid = {1};
        return 0;
    }
    std::vector<double> gy(d->node_y, d->node_y + d->n_nodes), gx(d->node_x, d->node_x + d->n_nodes);
    P.ygrid = gy;
    P.xgrid = gx;
    std::sort(P.ygrid.begin(), P.ygrid.end());
    P.ygrid.erase(std::unique(P.ygrid.begin(), P.ygrid.end()), P.ygrid.end());
    std::sort(P.xgrid.begin(), P.xgrid.end());
    P.xgrid.erase(std::unique(P.xgrid.begin(), P.xgrid.end()), P.xgrid.end());
    int ngy = (int)P.ygrid.size(), ngx = (int)P.xgrid.size();
    if (ngx < 2 || ngy < 2) return fail("`grid_y` and `grid_x` must each contain at least 2 unique values");
    if (ngx * ngy != d->n_nodes) return fail("node positions do not form a complete rectangular grid");
    P.index_grid.assign((size_t)ngy * ngx, 0);
    for (int i = 0; i < d->n_nodes; ++i) {
        int iy = (int)(std::lower_bound(P.ygrid.begin(), P.ygrid.end(), gy[i]) - P.ygrid.begin());
        int ix = (int)(std::lower_bound(P.xgrid.begin(), P.xgrid.end(), gx[i]) - P.xgrid.begin());
        int& slot = P.index_grid[(size_t)iy + (size_t)ix * ngy];
        if (slot != 0) return fail("duplicate node position");
        slot = i + 1;
    }
    return 0;
}

struct Corner {
    int idx;  // 1-based node id, 0 = inactive
    double w, dwdy, dwdx;
};

// _grid_corners_dw (gridded_psf.jl:176-208).
void grid_corners_dw(const Psf& P, double Y, double X, Corner c[4]) {
    if (P.n_nodes == 1) {
        c[0] = {1, 1.0, 0.0, 0.0};
        c[1] = c[2] = c[3] = {0, 0.0, 0.0, 0.0};
        return;
    }
    int ny = (int)P.ygrid.size(), nx = (int)P.xgrid.size();
    // searchsortedlast: index (1-based) of the last element <= v, 0 if none
    int iy = (int)(std::upper_bound(P.ygrid.begin(), P.ygrid.end(), Y) - P.ygrid.begin());
    int ix = (int)(std::upper_bound(P.xgrid.begin(), P.xgrid.end(), X) - P.xgrid.begin());
    iy = std::min(std::max(iy, 1), ny - 1);
    ix = std::min(std::max(ix, 1), nx - 1);
    double y0 = P.ygrid[iy - 1], y1 = P.ygrid[iy];
    double x0 = P.xgrid[ix - 1], x1 = P.xgrid[ix];
    double Yc = std::min(std::max(Y, y0), y1);
    double Xc = std::min(std::max(X, x0), x1);
    double cy = (Y > y0 && Y < y1) ? 1.0 : 0.0;
    double cx = (X > x0 && X < x1) ? 1.0 : 0.0;
    double invnorm = 1.0 / ((x1 - x0) * (y1 - y0));
    double dy1 = y1 - Yc, dy0 = Yc - y0, dx1 = x1 - Xc, dx0 = Xc - x0;
    auto ig = [&](int a, int b) { return P.index_grid[(size_t)(a - 1) + (size_t)(b - 1) * ny]; };
    c[0] = {ig(iy, ix), dx1 * dy1 * invnorm, -dx1 * cy * invnorm, -dy1 * cx * invnorm};
    c[1] = {ig(iy, ix + 1), dx0 * dy1 * invnorm, -dx0 * cy * invnorm, dy1 * cx * invnorm};
    c[2] = {ig(iy + 1, ix), dx1 * dy0 * invnorm, dx1 * cy * invnorm, -dy0 * cx * invnorm};
    c[3] = {ig(iy + 1, ix + 1), dx0 * dy0 * invnorm, dx0 * cy * invnorm, dy0 * cx * invnorm};
}

int nparams_kind(int kind) {
    switch (kind) {
        case CP_PSF_CIRCULAR_GAUSSIAN: return 5;
        case CP_PSF_CIRCULAR_MOFFAT: return 6;
        case CP_PSF_GRIDDED_IMAGE: return 4;
    }
    return -1;
}
int flux_slot(int kind) { return nparams_kind(kind) - 2; }
int bkg_slot(int kind) { return nparams_kind(kind) - 1; }

// evaluate (parametric_models.jl:77-84, :693-697; gridded_psf.jl:241-250 over
// empirical_models.jl:294-301).  prm is the getproperties-ordered vector.
double evaluate(const Psf& P, const double* prm, double py, double px) {
    if (P.kind == CP_PSF_CIRCULAR_GAUSSIAN) {
        double y = prm[0], x = prm[1], fwhm = prm[2], flux = prm[3], bkg = prm[4];
        double dx = px - x, dy = py - y;
        double sq = (dx * dx + dy * dy) / (fwhm * fwhm);
        double norm = -(M_PI * fwhm * fwhm / GAUSS_PRE);
        double amp = flux / norm;
        return std::fma(amp, std::exp(GAUSS_PRE * sq), bkg);
    }
    if (P.kind == CP_PSF_CIRCULAR_MOFFAT) {
        double y = prm[0], x = prm[1], a = prm[2], b = prm[3], flux = prm[4], bkg = prm[5];
        double r2 = (px - x) * (px - x) + (py - y) * (py - y);
        double amp = flux * (b - 1) / (M_PI * a * a);
        return std::fma(amp, std::pow(1 + r2 / (a * a), -b), bkg);
    }
    double Y = prm[0], X = prm[1], flux = prm[2], bkg = prm[3];
    Corner c[4];
    grid_corners_dw(P, Y, X, c);
    double s = 0;
    size_t nn = (size_t)P.gny * P.gnx;
    for (int k = 0; k < 4; ++k) {
        if (c[k].w == 0.0) continue;
        double u = P.os_x * (px - X) + P.ox;
        double v = P.os_y * (py - Y) + P.oy;
        double o[3];
        bicubic(P.data.data() + (size_t)(c[k].idx - 1) * nn, P.gny, P.gnx, v, u, P.fill, o);
        s = std::fma(c[k].w, o[0], s);
    }
    return std::fma(flux, s, bkg);
}

// evaluate_fg (parametric_models.jl:92-113, :704-724; gridded_psf.jl:252-271
// over empirical_models.jl:304-318).
void evaluate_fg(const Psf& P, const double* prm, double py, double px, double& f, double* G) {
    if (P.kind == CP_PSF_CIRCULAR_GAUSSIAN) {
        double y = prm[0], x = prm[1], fwhm = prm[2], flux = prm[3], bkg = prm[4];
        double dx = px - x, dy = py - y;
        double f2 = fwhm * fwhm;
        double sq = (dx * dx + dy * dy) / f2;
        double norm = -(M_PI * f2 / GAUSS_PRE);
        double amp = flux / norm;
        double g = std::exp(GAUSS_PRE * sq);
        double Ag = amp * g;
        f = std::fma(amp, g, bkg);
        double gf2 = GAUSS_PRE / f2;
        G[0] = -2 * Ag * gf2 * dy;
        G[1] = -2 * Ag * gf2 * dx;
        G[2] = -2 * Ag * (1 + GAUSS_PRE * sq) / fwhm;
        G[3] = g / norm;
        G[4] = 1.0;
        return;
    }
    if (P.kind == CP_PSF_CIRCULAR_MOFFAT) {
        double y = prm[0], x = prm[1], a = prm[2], b = prm[3], flux = prm[4], bkg = prm[5];
        double a2 = a * a;
        double dx = px - x, dy = py - y;
        double r2 = dx * dx + dy * dy;
        double u = 1 + r2 / a2;
        double profile = std::pow(u, -b);
        double norm = M_PI * a2 / (b - 1);
        double amp = flux / norm;
        double Ag = amp * profile;
        f = std::fma(amp, profile, bkg);
        G[0] = 2 * Ag * b * dy / (a2 * u);
        G[1] = 2 * Ag * b * dx / (a2 * u);
        G[2] = Ag * (-2 / a + 2 * b * r2 / (a * a2 * u));
        G[3] = Ag * (1 / (b - 1) - std::log(u));
        G[4] = profile / norm;
        G[5] = 1.0;
        return;
    }
    double Y = prm[0], X = prm[1], flux = prm[2], bkg = prm[3];
    Corner c[4];
    grid_corners_dw(P, Y, X, c);
    double s = 0, dsdy = 0, dsdx = 0;
    size_t nn = (size_t)P.gny * P.gnx;
    for (int k = 0; k < 4; ++k) {
        if (c[k].w == 0.0) continue;
        double u = P.os_x * (px - X) + P.ox;
        double v = P.os_y * (py - Y) + P.oy;
        double o[3];
        bicubic(P.data.data() + (size_t)(c[k].idx - 1) * nn, P.gny, P.gnx, v, u, P.fill, o);
        double phi = o[0];
        double g1 = -1.0 * P.os_y * o[1];  // unit-flux node gradient d/dY
        double g2 = -1.0 * P.os_x * o[2];  // d/dX
        s = std::fma(c[k].w, phi, s);
        dsdy = std::fma(c[k].dwdy, phi, std::fma(c[k].w, g1, dsdy));
        dsdx = std::fma(c[k].dwdx, phi, std::fma(c[k].w, g2, dsdx));
    }
    f = std::fma(flux, s, bkg);
    G[0] = flux * dsdy;
    G[1] = flux * dsdx;
    G[2] = s;
    G[3] = 1.0;
}

// Unit-flux curve of growth for :auto model radii (curve_of_growth.jl:180-211).
double cog_pixels(const Psf& P, double r);

double cog_unit(const Psf& P, double r) {
    if (P.kind == CP_PSF_CIRCULAR_GAUSSIAN) return 1 - std::exp(GAUSS_PRE * r * r / (P.fwhm * P.fwhm));
    if (P.kind == CP_PSF_CIRCULAR_MOFFAT)
        return 1 - std::pow(1 + r * r / (P.alpha * P.alpha), 1 - P.beta);
    return cog_pixels(P, r);   // no closed form: the generic aperture sum
}

// Exact area of the rectangle [x0, x1] x [y0, y1] inside the circle of radius
// r about the origin, as the integral over x of the covered vertical extent
// max(0, min(y1, h) - max(y0, -h)), h = sqrt(r^2 - x^2).  The extent's form
// changes only where h or -h crosses y0 or y1 (x = +-sqrt(r^2 - y^2)), so the
// integral is split there and each piece, whose form its midpoint decides, is
// integrated in closed form (int h dx = (x h + r^2 asin(x / r)) / 2).
double rect_circle_area(double x0, double y0, double x1, double y1, double r) {
    const double a = std::max(x0, -r), b = std::min(x1, r);
    if (!(b > a)) return 0.0;
    // x = 0 is a cut too: there h reaches r, so a bound at |y| = r touches
    // the circle without crossing it and no piece has it as its midpoint
    std::vector<double> cut = {a, b};
    if (a < 0 && b > 0) cut.push_back(0.0);
    for (double y : {y0, y1})
        if (std::fabs(y) < r) {
            const double s = std::sqrt((r - y) * (r + y));
            for (double c : {-s, s})
                if (c > a && c < b) cut.push_back(c);
        }
    std::sort(cut.begin(), cut.end());
    auto hh = [&](double x) { return std::sqrt(std::max(0.0, (r - x) * (r + x))); };
    auto Hi = [&](double x) { return 0.5 * (x * hh(x) + r * r * std::asin(std::min(1.0, std::max(-1.0, x / r)))); };
    double area = 0;
    for (size_t k = 0; k + 1 < cut.size(); ++k) {
        const double u = cut[k], v = cut[k + 1];
        if (!(v > u)) continue;
        const double m = 0.5 * (u + v), hm = hh(m);
        const bool top_is_h = hm < y1, bot_is_h = -hm > y0;
        const double top = top_is_h ? hm : y1, bot = bot_is_h ? -hm : y0;
        if (!(top > bot)) continue;
        const double It = top_is_h ? Hi(v) - Hi(u) : y1 * (v - u);
        const double Ib = bot_is_h ? -(Hi(v) - Hi(u)) : y0 * (v - u);
        area += It - Ib;
    }
    return area;
}

// aperture_weight(ap, i, j, ExactOverlap()) (apertures.jl:93-166) for the
// pixel centred at offset (dx, dy) from the aperture centre.
double pixel_overlap(double dx, double dy, double r) {
    const double d2 = dx * dx + dy * dy, dr = std::sqrt(2.0) / 2;
    if (d2 > (r + dr) * (r + dr)) return 0.0;
    if (r > dr && d2 < (r - dr) * (r - dr)) return 1.0;
    return rect_circle_area(dx - 0.5, dy - 0.5, dx + 0.5, dy + 0.5, r);
}

// curve_of_growth(model, radii) generic method (curve_of_growth.jl:143-176)
// for the unit model at (0, 0) over bounding_axes (apertures.jl:52-65).
double cog_pixels(const Psf& P, double r) {
    double prm[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (P.kind == CP_PSF_CIRCULAR_GAUSSIAN) {
        prm[2] = P.fwhm;
        prm[3] = 1;
    } else if (P.kind == CP_PSF_CIRCULAR_MOFFAT) {
        prm[2] = P.alpha;
        prm[3] = P.beta;
        prm[4] = 1;
    } else {
        prm[2] = 1;
    }
    const double lo = -r - 0.5, hi = r - 0.5;
    const long i0 = (long)(std::ceil(lo) == lo ? std::ceil(lo) + 1 : std::ceil(lo)), i1 = (long)std::ceil(hi);
    double s = 0;
    for (long j = i0; j <= i1; ++j)
        for (long i = i0; i <= i1; ++i) {
            const double w = pixel_overlap((double)j, (double)i, r);
            if (w > 0) s += w * evaluate(P, prm, (double)i, (double)j);
        }
    return s;
}

// Julia default quantile (type 7) of an unsorted sample.
double quantile7(std::vector<double> v, double p) {
    std::sort(v.begin(), v.end());
    size_t n = v.size();
    double h = (n - 1) * p;
    size_t lo = (size_t)std::floor(h);
    if (lo + 1 >= n) return v[n - 1];
    return v[lo] + (h - lo) * (v[lo + 1] - v[lo]);
}

// Julia round(Int, x): round half to even; false when not representable.
bool round_even_int(double v, int64_t& out) {
    if (!std::isfinite(v)) return false;
    double r = std::nearbyint(v);  // default rounding mode: nearest-even
    if (std::fabs(r) > 1073741824.0) return false;
    out = (int64_t)r;
    return true;
}

// ---------------------------------------------------------------------------
// damped least-squares inner solvers on an abstract operator
// ---------------------------------------------------------------------------

struct Op {
    int64_t m, n;
    std::function<void(const double*, double*)> J;   // out(m) = J v(n)
    std::function<void(const double*, double*)> JT;  // out(n) = J^T u(m)
};

double dot(const double* a, const double* b, int64_t n) {
    double s = 0;
    for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}
double nrm2(const double* a, int64_t n) { return std::sqrt(dot(a, a, n)); }
double sgn(double v) { return v > 0 ? 1.0 : (v < 0 ? -1.0 : 0.0); }

// Symmetric Givens rotation (c, s, rho) with [c s; s -c][a; b] = [rho; 0].
void sym_givens(double a, double b, double& c, double& s, double& rho) {
    if (b == 0) {
        c = (a == 0) ? 1.0 : sgn(a);
        s = 0;
        rho = std::fabs(a);
    } else if (a == 0) {
        c = 0;
        s = sgn(b);
        rho = std::fabs(b);
    } else if (std::fabs(b) > std::fabs(a)) {
        double t = a / b;
        double tau = sgn(b) * std::sqrt(1 + t * t);
        s = 1 / tau;
        c = s * t;
        rho = b * tau;
    } else {
        double t = b / a;
        double tau = sgn(a) * std::sqrt(1 + t * t);
        c = 1 / tau;
        s = c * t;
        rho = a * tau;
    }
}

// LSQR, Paige & Saunders, ACM TOMS 8(1):43-71 (1982), with damping lam and the
// Krylov.jl-style stopping tests (atol on ||A'r||, btol/axtol on ||r||,
// conlim, a 5-deep forward-error window).  Called as in
// psf_photometry_simultaneous.jl:1221-1224 with x0 = 0.
int lsqr(const Op& A, const double* b, double lam, int itmax, double atol, double btol,
         double* x) {
    const int64_t m = A.m, n = A.n;
    const double axtol = std::sqrt(EPS64), etol = std::sqrt(EPS64);
    const double ctol = 1.0 / (1.0 / std::sqrt(EPS64));
    const int window = 5;
    std::vector<double> u(b, b + m), v(n), w(n), Av(m), Atu(n), err_vec(window, 0.0);
    std::fill(x, x + n, 0.0);
    double beta1 = nrm2(u.data(), m);
    if (beta1 == 0) return 0;
    for (auto& t : u) t /= beta1;
    A.JT(u.data(), v.data());
    double Anorm2 = dot(v.data(), v.data(), n);
    double alpha = std::sqrt(Anorm2);
    if (alpha == 0) return 0;
    for (auto& t : v) t /= alpha;
    w = v;
    double phibar = beta1, rhobar = alpha;
    double xNorm2 = 0, dNorm2 = 0, c2 = -1, s2 = 0, z = 0, res2 = 0, xENorm2 = 0, err_lbnd = 0;
    double lam2 = lam * lam;
    int iter = 0;
    bool done = false;
    while (!done) {
        ++iter;
        A.J(v.data(), Av.data());
        for (int64_t i = 0; i < m; ++i) u[i] = Av[i] - alpha * u[i];
        double beta = nrm2(u.data(), m);
        if (beta != 0) {
            for (auto& t : u) t /= beta;
            Anorm2 += alpha * alpha + beta * beta;
            if (lam > 0) Anorm2 += lam2;
            A.JT(u.data(), Atu.data());
            for (int64_t i = 0; i < n; ++i) v[i] = Atu[i] - beta * v[i];
            alpha = nrm2(v.data(), n);
            if (alpha != 0)
                for (auto& t : v) t /= alpha;
        }
        double c1, s1, rhobar1;
        sym_givens(rhobar, lam, c1, s1, rhobar1);
        double psi = s1 * phibar;
        phibar = c1 * phibar;
        double c, s, rho;
        sym_givens(rhobar1, beta, c, s, rho);
        double theta = s * alpha;
        rhobar = -c * alpha;
        double phi = c * phibar;
        phibar = s * phibar;
        xENorm2 += phi * phi;
        err_vec[iter % window] = phi;
        if (iter >= window) err_lbnd = nrm2(err_vec.data(), window);
        double tau = s * phi;
        dNorm2 += dot(w.data(), w.data(), n) / (rho * rho);
        double sig = phi / rho, th = theta / rho;
        for (int64_t i = 0; i < n; ++i) {
            x[i] += sig * w[i];
            w[i] = v[i] - th * w[i];
        }
        double delta = s2 * rho, gambar = -c2 * rho, rhs = phi - delta * z;
        double zbar = rhs / gambar;
        double xNorm = std::sqrt(xNorm2 + zbar * zbar);
        double gamma;
        sym_givens(gambar, theta, c2, s2, gamma);
        z = rhs / gamma;
        xNorm2 += z * z;
        double Anorm = std::sqrt(Anorm2);
        double Acond = Anorm * std::sqrt(dNorm2);
        double res1 = phibar * phibar;
        res2 += psi * psi;
        double rNorm = std::sqrt(res1 + res2);
        double ArNorm = alpha * std::fabs(tau);
        double test1 = rNorm / beta1;
        double test2 = ArNorm / (Anorm * rNorm);
        double test3 = 1 / Acond;
        double t1 = test1 / (1 + Anorm * xNorm / beta1);
        double rNormtol = btol + axtol * Anorm * xNorm / beta1;
        bool ill = (1 + test3 <= 1) || (test3 <= ctol);
        bool solved = (1 + test2 <= 1) || (test2 <= axtol) || (ArNorm <= atol) ||
                      (1 + t1 <= 1) || (test1 <= rNormtol) ||
                      (iter >= window && err_lbnd <= etol * std::sqrt(xENorm2));
        done = solved || ill || iter >= itmax;
    }
    return iter;
}

// LSMR, Fong & Saunders, SIAM J. Sci. Comput. 33(5):2950-2971 (2011), with
// damping lam and the stopping tests of the Krylov.jl implementation the
// reference calls (solver = :lsmr, psf_photometry_simultaneous.jl:1164,
// 1221-1224): atol on ||A'r||, btol / axtol on ||r||, conlim, and the 5-deep
// forward-error window on zeta.  ||x|| is the norm of the current iterate.
int lsmr(const Op& A, const double* b, double lam, int itmax, double atol, double btol, double* x) {
    const int64_t m = A.m, n = A.n;
    const double axtol = std::sqrt(EPS64), etol = std::sqrt(EPS64);
    const double ctol = 1.0 / (1.0 / std::sqrt(EPS64));
    const int window = 5;
    std::vector<double> u(b, b + m), v(n), h(n), hbar(n, 0.0), Av(m), Atu(n), err_vec(window, 0.0);
    std::fill(x, x + n, 0.0);
    double beta1 = nrm2(u.data(), m);
    if (beta1 == 0) return 0;
    for (auto& t : u) t /= beta1;
    A.JT(u.data(), v.data());
    double alpha = nrm2(v.data(), n);
    if (alpha == 0) return 0;
    for (auto& t : v) t /= alpha;
    h = v;
    double zetabar = alpha * beta1, alphabar = alpha, rho = 1, rhobar = 1, cbar = 1, sbar = 0;
    double betadd = beta1, betad = 0, rhodold = 1, tautildeold = 0, thetatilde = 0, zeta = 0, d = 0;
    double Anorm2 = alpha * alpha, maxrbar = 0, minrbar = 1e100;
    double xENorm2 = 0, err_lbnd = 0;
    int iter = 0;
    bool done = false;
    while (!done) {
        ++iter;
        // Golub-Kahan step: beta u = A v - alpha u, alpha v = A' u - beta v
        A.J(v.data(), Av.data());
        for (int64_t i = 0; i < m; ++i) u[i] = Av[i] - alpha * u[i];
        double beta = nrm2(u.data(), m);
        if (beta != 0) {
            for (auto& t : u) t /= beta;
            A.JT(u.data(), Atu.data());
            for (int64_t i = 0; i < n; ++i) v[i] = Atu[i] - beta * v[i];
            alpha = nrm2(v.data(), n);
            if (alpha != 0)
                for (auto& t : v) t /= alpha;
        }
        // the damping rotation, then the two QR rotations
        double chat, shat, alphahat;
        sym_givens(alphabar, lam, chat, shat, alphahat);
        const double rhoold = rho;
        double c, s;
        sym_givens(alphahat, beta, c, s, rho);
        const double thetanew = s * alpha;
        alphabar = c * alpha;
        const double rhobarold = rhobar, zetaold = zeta;
        const double thetabar = sbar * rho, rhotemp = cbar * rho;
        sym_givens(rhotemp, thetanew, cbar, sbar, rhobar);
        zeta = cbar * zetabar;
        zetabar = -sbar * zetabar;
        xENorm2 += zeta * zeta;
        err_vec[iter % window] = zeta;
        if (iter >= window) err_lbnd = nrm2(err_vec.data(), window);
        // h, hbar and x
        const double delta = thetabar * rho / (rhoold * rhobarold);
        const double sigma = zeta / (rho * rhobar), eta = thetanew / rho;
        for (int64_t i = 0; i < n; ++i) {
            hbar[i] = h[i] - delta * hbar[i];
            x[i] += sigma * hbar[i];
            h[i] = v[i] - eta * h[i];
        }
        // ||r|| estimate
        const double betaacute = chat * betadd, betacheck = -shat * betadd;
        const double betahat = c * betaacute;
        betadd = -s * betaacute;
        const double thetatildeold = thetatilde;
        double ctildeold, stildeold, rhotildeold;
        sym_givens(rhodold, thetabar, ctildeold, stildeold, rhotildeold);
        thetatilde = stildeold * rhobar;
        rhodold = ctildeold * rhobar;
        betad = -stildeold * betad + ctildeold * betahat;
        tautildeold = (zetaold - thetatildeold * tautildeold) / rhotildeold;
        const double taud = (zeta - thetatilde * tautildeold) / rhodold;
        d += betacheck * betacheck;
        const double rNorm = std::sqrt(d + (betad - taud) * (betad - taud) + betadd * betadd);
        // ||A|| and cond(A)
        Anorm2 += beta * beta;
        const double Anorm = std::sqrt(Anorm2);
        Anorm2 += alpha * alpha;
        maxrbar = std::max(maxrbar, rhobarold);
        if (iter > 1) minrbar = std::min(minrbar, rhobarold);
        const double Acond = std::max(maxrbar, rhotemp) / std::min(minrbar, rhotemp);
        // tests
        const double ArNorm = std::fabs(zetabar);
        const double xNorm = nrm2(x, n);
        const double test1 = rNorm / beta1;
        const double test2 = ArNorm / (Anorm * rNorm);
        const double test3 = 1 / Acond;
        const double t1 = test1 / (1 + Anorm * xNorm / beta1);
        const double rNormtol = btol + axtol * Anorm * xNorm / beta1;
        const bool ill = (1 + test3 <= 1) || (test3 <= ctol);
        const bool solved = (1 + test2 <= 1) || (test2 <= axtol) || (ArNorm <= atol) || (1 + t1 <= 1) ||
                            (test1 <= rNormtol) || (iter >= window && err_lbnd <= etol * std::sqrt(xENorm2));
        done = solved || ill || iter >= itmax;
    }
    return iter;
}

// CG (Hestenes-Stiefel) on the damped normal equations (J'J + lam^2 I) x = J'b:
// the search direction p, its image q = J p, z = J'q + lam^2 p = N p, the
// normal-equation residual s updated by s <- s - a z (the recurrence the
// device runs), gamma = |s|^2 computed from s itself.  The build's stopping
// rule (CG has no reference counterpart): at least one step, then stop when
// ||s_k|| <= atol * ||s_0|| or after itmax steps; a step with p'Np <= 0 ends
// the solve before it is taken.
int cgls(const Op& A, const double* b, double lam, int itmax, double atol, double /*btol*/,
         double* x) {
    const int64_t m = A.m, n = A.n;
    std::vector<double> s(n), p(n), q(m), z(n);
    std::fill(x, x + n, 0.0);
    A.JT(b, s.data());
    p = s;
    double gamma = dot(s.data(), s.data(), n);
    const double stop = atol * std::sqrt(gamma);
    double lam2 = lam * lam;
    int iter = 0;
    if (gamma == 0) return 0;
    while (iter < itmax) {
        ++iter;
        A.J(p.data(), q.data());
        A.JT(q.data(), z.data());
        for (int64_t i = 0; i < n; ++i) z[i] += lam2 * p[i];
        double delta = dot(q.data(), q.data(), m) + lam2 * dot(p.data(), p.data(), n);
        if (!(delta > 0)) break;
        double a = gamma / delta;
        for (int64_t i = 0; i < n; ++i) {
            x[i] += a * p[i];
            s[i] -= a * z[i];
        }
        double gnew = dot(s.data(), s.data(), n);
        if (std::sqrt(gnew) <= stop) break;
        double bet = gnew / gamma;
        for (int64_t i = 0; i < n; ++i) p[i] = s[i] + bet * p[i];
        gamma = gnew;
    }
    return iter;
}

// ---------------------------------------------------------------------------
// stamp operators (psf_photometry_simultaneous.jl:55-133)
// ---------------------------------------------------------------------------

void apply_JT_impl(double* z, const double* V, const int32_t* pixels, int p, int S2, int na,
                   const double* u, const uint8_t* live) {
    std::fill(z, z + (size_t)na * p, 0.0);
    for (int a = 0; a < na; ++a) {
        if (live && !live[a]) continue;
        for (int k = 0; k < p; ++k) {
            double acc = 0;
            for (int m = 0; m < S2; ++m) {
                int32_t fi = pixels[(size_t)a * S2 + m];
                acc += V[((size_t)a * S2 + m) * p + k] * u[(fi == 0 ? 1 : fi) - 1];
            }
            z[(size_t)a * p + k] = acc;
        }
    }
}

void apply_J_impl(double* y, int64_t npix, const double* V, const int32_t* pixels, int p, int S2,
                  int na, const double* v, const uint8_t* live) {
    std::fill(y, y + npix, 0.0);
    for (int a = 0; a < na; ++a) {
        if (live && !live[a]) continue;
        for (int m = 0; m < S2; ++m) {
            int32_t fi = pixels[(size_t)a * S2 + m];
            if (fi == 0) continue;
            double acc = 0;
            for (int k = 0; k < p; ++k) acc += V[((size_t)a * S2 + m) * p + k] * v[(size_t)a * p + k];
            y[fi - 1] += acc;
        }
    }
}

// ---------------------------------------------------------------------------
// small dense linear algebra for covariance (levenberg_marquardt.jl:366-388)
// ---------------------------------------------------------------------------

// Cholesky of a symmetric p x p (upper triangle used); false if not positive definite.
bool cholesky(const double* A, int p, double* L) {
    std::fill(L, L + p * p, 0.0);
    for (int j = 0; j < p; ++j) {
        double s = A[j * p + j];
        for (int k = 0; k < j; ++k) s -= L[j * p + k] * L[j * p + k];
        if (!(s > 0)) return false;
        double d = std::sqrt(s);
        L[j * p + j] = d;
        for (int i = j + 1; i < p; ++i) {
            double t = A[std::min(i, j) * p + std::max(i, j)];  // upper triangle (row<=col), row-major view
            for (int k = 0; k < j; ++k) t -= L[i * p + k] * L[j * p + k];
            L[i * p + j] = t / d;
        }
    }
    return true;
}

void chol_solve(const double* L, int p, const double* b, double* x) {
    std::vector<double> y(p);
    for (int i = 0; i < p; ++i) {
        double s = b[i];
        for (int k = 0; k < i; ++k) s -= L[i * p + k] * y[k];
        y[i] = s / L[i * p + i];
    }
    for (int i = p - 1; i >= 0; --i) {
        double s = y[i];
        for (int k = i + 1; k < p; ++k) s -= L[k * p + i] * x[k];
        x[i] = s / L[i * p + i];
    }
}

// Jacobi eigen-decomposition of a symmetric matrix (for pinv).
void jacobi_eig(const double* A0, int p, double* evals, double* evecs) {
    std::vector<double> A(A0, A0 + p * p);
    for (int i = 0; i < p; ++i)
        for (int j = 0; j < p; ++j) evecs[i * p + j] = (i == j);
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0;
        for (int i = 0; i < p; ++i)
            for (int j = i + 1; j < p; ++j) off += A[i * p + j] * A[i * p + j];
        if (off < 1e-300) break;
        for (int i = 0; i < p; ++i)
            for (int j = i + 1; j < p; ++j) {
                double aij = A[i * p + j];
                if (aij == 0) continue;
                double th = (A[j * p + j] - A[i * p + i]) / (2 * aij);
                double t = sgn(th) / (std::fabs(th) + std::sqrt(th * th + 1));
                if (th == 0) t = 1;
                double c = 1 / std::sqrt(t * t + 1), s = t * c;
                for (int k = 0; k < p; ++k) {
                    double aki = A[k * p + i], akj = A[k * p + j];
                    A[k * p + i] = c * aki - s * akj;
                    A[k * p + j] = s * aki + c * akj;
                }
                for (int k = 0; k < p; ++k) {
                    double aik = A[i * p + k], ajk = A[j * p + k];
                    A[i * p + k] = c * aik - s * ajk;
                    A[j * p + k] = s * aik + c * ajk;
                }
                for (int k = 0; k < p; ++k) {
                    double vki = evecs[k * p + i], vkj = evecs[k * p + j];
                    evecs[k * p + i] = c * vki - s * vkj;
                    evecs[k * p + j] = s * vki + c * vkj;
                }
            }
    }
    for (int i = 0; i < p; ++i) evals[i] = A[i * p + i];
}

// inv via Cholesky, else pinv (rtol = eps * p relative to the largest |eigenvalue|).
void sym_inverse(const double* A, int p, double* out) {
    std::vector<double> L(p * p), e(p), col(p);
    if (cholesky(A, p, L.data())) {
        for (int j = 0; j < p; ++j) {
            std::vector<double> ej(p, 0.0);
            ej[j] = 1;
            chol_solve(L.data(), p, ej.data(), col.data());
            for (int i = 0; i < p; ++i) out[i * p + j] = col[i];
        }
        return;
    }
    std::vector<double> Vv(p * p);
    jacobi_eig(A, p, e.data(), Vv.data());
    double smax = 0;
    for (int i = 0; i < p; ++i) smax = std::max(smax, std::fabs(e[i]));
    double tol = EPS64 * p * smax;
    for (int i = 0; i < p; ++i)
        for (int j = 0; j < p; ++j) {
            double s = 0;
            for (int k = 0; k < p; ++k)
                if (std::fabs(e[k]) > tol) s += Vv[i * p + k] * Vv[j * p + k] / e[k];
            out[i * p + j] = s;
        }
}

// ---------------------------------------------------------------------------
// lm_irls (levenberg_marquardt.jl:431-680), no IRLS reweighting.
// accum(A, b, x) fills the p x p normal matrix and gradient, returns cost.
// ---------------------------------------------------------------------------

using Accum = std::function<double(double* A, double* b, const double* x)>;

void damp(double* A, int p, int damping, double min_diag, double lam) {
    for (int i = 0; i < p; ++i) {
        if (damping == CP_DAMP_MARQUARDT)
            A[i * p + i] += lam * std::max(A[i * p + i], min_diag);
        else if (damping == CP_DAMP_LEVENBERG)
            A[i * p + i] += lam;
    }
}

int lm_irls(const Accum& accum, int p, int nobs, bool have_weights, std::vector<double>& x,
            const cp_lm_opts& o, cp_lm_stat& st, double* cov_out) {
    if (p <= 0) return fail("all parameters are fixed; nothing to fit");
    if (nobs <= 0) return fail("number of observations must be positive");
    int dof = nobs - p;
    if (dof <= 0) return fail("degrees of freedom must be positive");
    double lam = o.lambda_init;
    std::vector<double> A(p * p), b(p), Ac(p * p), bc(p), Ad(p * p), L(p * p), d(p), xc(p), D(p);
    double cost = accum(A.data(), b.data(), x.data());
    double cost_init = cost;
    const double d_floor = 1e-6;
    for (int i = 0; i < p; ++i) D[i] = std::sqrt(std::max(A[i * p + i], d_floor));
    bool xconv = false, fconv = false, gconv = false, conv = false;
    int iter = 0;
    while (iter < o.max_iter) {
        ++iter;
        double gcos = 0;
        for (int i = 0; i < p; ++i)
            gcos = std::max(gcos, std::fabs(b[i]) / (std::sqrt(A[i * p + i] * cost) + EPS64));
        if (gcos <= o.g_tol) {
            gconv = conv = true;
            break;
        }
        Ad = A;
        damp(Ad.data(), p, o.damping, o.min_diagonal, lam);
        if (!cholesky(Ad.data(), p, L.data())) {
            lam = std::min(lam * o.lambda_up, o.lambda_max);
            continue;
        }
        chol_solve(L.data(), p, b.data(), d.data());
        for (int i = 0; i < p; ++i) d[i] = -d[i];
        double dAd = 0, db = 0;
        for (int i = 0; i < p; ++i) {
            db += d[i] * b[i];
            for (int j = 0; j < p; ++j) dAd += d[i] * A[i * p + j] * d[j];
        }
        double prered = -2 * db - dAd;
        for (int i = 0; i < p; ++i) xc[i] = x[i] + d[i];
        double cost_c = accum(Ac.data(), bc.data(), xc.data());
        bool accepted = cost_c < cost;
        double dcost = cost - cost_c;
        double ratio = prered > 0 ? dcost / prered : 0.0;
        fconv = std::fabs(dcost) <= o.f_tol * (std::fabs(cost) + o.f_tol) &&
                prered <= o.f_tol * (std::fabs(cost) + o.f_tol) && 0.5 * ratio <= 1;
        double nd = 0, nx = 0;
        for (int i = 0; i < p; ++i) {
            nd += (D[i] * d[i]) * (D[i] * d[i]);
            nx += (D[i] * x[i]) * (D[i] * x[i]);
        }
        xconv = std::sqrt(nd) <= o.x_tol * (std::sqrt(nx) + o.x_tol) && (accepted || fconv);
        if (accepted) {
            cost = cost_c;
            lam = std::max(lam / o.lambda_down, o.lambda_min);
            x = xc;
            A = Ac;
            b = bc;
            for (int i = 0; i < p; ++i) D[i] = std::max(D[i], std::sqrt(A[i * p + i]));
            if (xconv || fconv) {
                conv = true;
                break;
            }
        } else {
            if (xconv || fconv) {
                conv = true;
                break;
            }
            lam = std::min(lam * o.lambda_up, o.lambda_max);
        }
    }
    int ce = o.covariance;
    if (ce == CP_COV_DEFAULT) ce = have_weights ? CP_COV_KNOWN : CP_COV_REWEIGHTED;
    if (cov_out && ce != CP_COV_NONE) {
        sym_inverse(A.data(), p, cov_out);
        if (ce == CP_COV_REWEIGHTED)
            for (int i = 0; i < p * p; ++i) cov_out[i] *= cost / dof;
    }
    st.minimum = cost;
    st.cost_init = cost_init;
    st.converged = conv;
    st.x_converged = xconv;
    st.f_converged = fconv;
    st.g_converged = gconv;
    st.iterations = iter;
    st.lambda_final = lam;
    st.chisq = cost / dof;
    st.status = 0;
    return 0;
}

}  // namespace

// ===========================================================================
// extern "C" surface
// ===========================================================================

extern "C" {

const char* or_last_error(void) { return g_err.c_str(); }

int or_nparams(int32_t kind) { return nparams_kind(kind); }

int or_evaluate_fg(const cp_psf_desc* psf, const double* prm, double py, double px, double* f,
                   double* G) {
    Psf P;
    if (make_psf(psf, P)) return CP_EARG;
    evaluate_fg(P, prm, py, px, *f, G);
    return 0;
}

double or_evaluate(const cp_psf_desc* psf, const double* prm, double py, double px) {
    Psf P;
    if (make_psf(psf, P)) return std::numeric_limits<double>::quiet_NaN();
    return evaluate(P, prm, py, px);
}

void or_bicubic(const double* data, int32_t ny, int32_t nx, double y, double x, double fill,
                double* out3) {
    bicubic(data, ny, nx, y, x, fill, out3);
}

int or_grid_corners_dw(const cp_psf_desc* psf, double Y, double X, int32_t* idx4, double* w4,
                       double* dwdy4, double* dwdx4) {
    Psf P;
    if (make_psf(psf, P)) return CP_EARG;
    Corner c[4];
    grid_corners_dw(P, Y, X, c);
    for (int k = 0; k < 4; ++k) {
        idx4[k] = c[k].idx;
        w4[k] = c[k].w;
        dwdy4[k] = c[k].dwdy;
        dwdx4[k] = c[k].dwdx;
    }
    return 0;
}

double or_curve_of_growth(const cp_psf_desc* psf, double r) {
    Psf P;
    if (make_psf(psf, P)) return std::numeric_limits<double>::quiet_NaN();
    return cog_unit(P, r);
}

// The generic aperture-sum curve of growth for any model (checks against the
// analytic forms) and the exact pixel overlap.
double or_curve_of_growth_pixels(const cp_psf_desc* psf, double r) {
    Psf P;
    if (make_psf(psf, P) || !(r > 0)) return std::numeric_limits<double>::quiet_NaN();
    return cog_pixels(P, r);
}

double or_pixel_overlap(double dx, double dy, double r) { return pixel_overlap(dx, dy, r); }

// _morton2d (psf_photometry_simultaneous.jl:163-174).
uint64_t or_morton2d(int64_t y, int64_t x) {
    auto spread = [](uint64_t v) {
        v &= 0x00000000ffffffffULL;
        v = (v | (v << 16)) & 0x0000ffff0000ffffULL;
        v = (v | (v << 8)) & 0x00ff00ff00ff00ffULL;
        v = (v | (v << 4)) & 0x0f0f0f0f0f0f0f0fULL;
        v = (v | (v << 2)) & 0x3333333333333333ULL;
        v = (v | (v << 1)) & 0x5555555555555555ULL;
        return v;
    };
    return spread((uint64_t)y) | (spread((uint64_t)x) << 1);
}

// _ring_offsets (psf_photometry_simultaneous.jl:181-190).
int or_ring_offsets(int32_t R, int32_t* dy, int32_t* dx) {
    int k = 0;
    for (int r = 0; r <= R; ++r)
        for (int a = -r; a <= r; ++a)
            for (int b = -r; b <= r; ++b) {
                if (std::max(std::abs(a), std::abs(b)) != r) continue;
                dy[k] = a;
                dx[k] = b;
                ++k;
            }
    return k;
}

// _model_radii (psf_photometry_simultaneous.jl:209-243).  model_rad > 0 is the
// scalar path; model_rad <= 0 is :auto.
int or_model_radii(const cp_psf_desc* psf, double model_rad, double nsigma, int32_t R_fit,
                   int32_t R_cap, const double* w, int64_t npix, const double* flux_init,
                   int32_t n, int32_t* out) {
    if (model_rad > 0) {
        int64_t r = (int64_t)std::ceil(model_rad);
        r = std::min<int64_t>(std::max<int64_t>(r, R_fit), R_cap);
        for (int j = 0; j < n; ++j) out[j] = (int32_t)r;
        return 0;
    }
    Psf P;
    if (make_psf(psf, P)) return CP_EARG;
    int64_t step = std::max<int64_t>(1, npix / 65536);
    std::vector<double> ws;
    for (int64_t i = 0; i < npix; i += step)
        if (w[i] > 0 && std::isfinite(w[i])) ws.push_back(w[i]);
    if (ws.empty()) return fail("no positive finite weights for the :auto noise estimate");
    double wv = quantile7(ws, 0.84);
    double sigma_bg = wv > 0 ? std::sqrt(1 / wv) : 1.0;
    std::vector<double> ee(R_cap), sb(R_cap);
    double last = cog_unit(P, (double)R_cap);
    for (int r = 1; r <= R_cap; ++r) ee[r - 1] = cog_unit(P, (double)r) / last;
    double prev = 0;
    for (int r = 1; r <= R_cap; ++r) {
        sb[r - 1] = std::max((ee[r - 1] - prev) / (M_PI * (double)(r * r - (r - 1) * (r - 1))), 0.0);
        prev = ee[r - 1];
    }
    double thr = nsigma * sigma_bg;
    for (int j = 0; j < n; ++j) {
        double F = std::max(flux_init[j], 0.0);
        int r = R_fit;
        while (r < R_cap && F * sb[r - 1] > thr) ++r;
        out[j] = r;
    }
    return 0;
}

// _build_stamps! + _touched_pixels (psf_photometry_simultaneous.jl:245-335).
// pixels is written for the active stars in Morton order (S2 per star).
int or_build_stamps(const double* w, int32_t ny, int32_t nx, const double* y0, const double* x0,
                    const double* flux0, int32_t n_stars, double fit_rad, int32_t* anchor_y,
                    int32_t* anchor_x, int32_t* active, int32_t* n_active, int32_t* pixels,
                    uint8_t* union_mask) {
    int R = (int)std::ceil(fit_rad);
    int S = 2 * R + 1, S2 = S * S;
    std::vector<int> dyo(S2), dxo(S2);
    int m = 0;
    for (int dx = -R; dx <= R; ++dx)
        for (int dy = -R; dy <= R; ++dy) {
            dyo[m] = dy;
            dxo[m] = dx;
            ++m;
        }
    std::vector<int32_t> pix((size_t)S2 * n_stars, 0);
    std::vector<char> ok(n_stars, 0);
    for (int i = 0; i < n_stars; ++i) {
        int64_t ay, ax;
        bool fy = round_even_int(y0[i], ay), fx = round_even_int(x0[i], ax);
        anchor_y[i] = fy ? (int32_t)ay : INT32_MIN;
        anchor_x[i] = fx ? (int32_t)ax : INT32_MIN;
        if (!fy || !fx) continue;
        bool any = false;
        for (int k = 0; k < S2; ++k) {
            int64_t gy = ay + dyo[k], gx = ax + dxo[k];
            if (gy >= 1 && gy <= ny && gx >= 1 && gx <= nx) {
                int64_t fi = gy + (gx - 1) * (int64_t)ny;
                if ((w ? w[fi - 1] : 1.0) > 0) {
                    pix[(size_t)i * S2 + k] = (int32_t)fi;
                    any = true;
                }
            }
        }
        // active: >= 1 usable pixel and finite y/x/flux; the Morton key of a
        // negative anchor is undefined in the reference (it throws), so such a
        // star is excluded here as a failure.
        ok[i] = any && std::isfinite(y0[i]) && std::isfinite(x0[i]) && std::isfinite(flux0[i]) &&
                ay >= 0 && ax >= 0;
    }
    std::vector<int32_t> act;
    for (int i = 0; i < n_stars; ++i)
        if (ok[i]) act.push_back(i);
    std::vector<uint64_t> key(n_stars, 0);
    for (int i : act) key[i] = or_morton2d(anchor_y[i], anchor_x[i]);
    std::stable_sort(act.begin(), act.end(), [&](int a, int b) { return key[a] < key[b]; });
    *n_active = (int32_t)act.size();
    for (size_t j = 0; j < act.size(); ++j) {
        active[j] = act[j];
        if (pixels) std::memcpy(pixels + j * S2, pix.data() + (size_t)act[j] * S2, sizeof(int32_t) * S2);
    }
    if (union_mask) {
        std::memset(union_mask, 0, (size_t)ny * nx);
        for (int i : act)
            for (int k = 0; k < S2; ++k) {
                int32_t fi = pix[(size_t)i * S2 + k];
                if (fi) union_mask[fi - 1] = 1;
            }
    }
    return 0;
}

void or_apply_J(double* y, int64_t npix, const double* V, const int32_t* pixels, int32_t p,
                int32_t S2, int32_t n_active, const double* v, const uint8_t* live) {
    apply_J_impl(y, npix, V, pixels, p, S2, n_active, v, live);
}

void or_apply_JT(double* z, const double* V, const int32_t* pixels, int32_t p, int32_t S2,
                 int32_t n_active, const double* u, const uint8_t* live) {
    apply_JT_impl(z, V, pixels, p, S2, n_active, u, live);
}

int or_inner_solve(int32_t solver, const double* V, const int32_t* pixels, int32_t p, int32_t S2,
                   int32_t n_active, int64_t npix, const uint8_t* live, const double* b,
                   double lam, int32_t itmax, double atol, double btol, double* x_out) {
    Op op;
    op.m = npix;
    op.n = (int64_t)p * n_active;
    op.J = [&](const double* v, double* out) { apply_J_impl(out, npix, V, pixels, p, S2, n_active, v, live); };
    op.JT = [&](const double* u, double* out) { apply_JT_impl(out, V, pixels, p, S2, n_active, u, live); };
    if (solver == CP_SOLVER_CGLS) return cgls(op, b, lam, itmax, atol, btol, x_out);
    if (solver == CP_SOLVER_LSMR) return lsmr(op, b, lam, itmax, atol, btol, x_out);
    return lsqr(op, b, lam, itmax, atol, btol, x_out);
}

}  // extern "C"

namespace {

// ---------------------------------------------------------------------------
// post-fit diagnostics (src/photometry/psf_photometry_diagnostics.jl)
// ---------------------------------------------------------------------------

// Mixture-of-Gaussians exponential disk (psf_photometry_diagnostics.jl:14-15).
const double EXPGAL_AMP[6] = {1.99485977e-4, 2.61612679e-3, 1.89726655e-2,
                              1.00186544e-1, 3.68534484e-1, 5.09490694e-1};
const double EXPGAL_VAR[6] = {1.20078965e-3, 8.84526493e-3, 3.91463084e-2,
                              1.39976817e-1, 4.60962500e-1, 1.50159566e0};

// 64-point Gauss-Legendre nodes / weights on [-1, 1] (Newton on P_64).
void gauss_legendre64(double* x, double* wgt) {
    const int n = 64;
    for (int i = 0; i < n; ++i) {
        double z = std::cos(M_PI * (i + 0.75) / (n + 0.5)), dp = 0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1, p1 = 0;
            for (int k = 1; k <= n; ++k) {
                const double p2 = p1;
                p1 = p0;
                p0 = ((2 * k - 1) * z * p1 - (k - 1) * p2) / k;
            }
            dp = n * (z * p0 - p1) / (z * z - 1);
            const double dz = p0 / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-16) break;
        }
        x[i] = z;
        wgt[i] = 2 / ((1 - z * z) * dp * dp);
    }
}

// I(m; v) of _exp_disk_kernel_bandlimited (:78-101): the integral
// int_{-1/2}^{1/2} exp(-a k^2) cos(2 pi k m) dk (the reference's closed form
// through the complex error function), by Gauss-Legendre on 8 panels.
double band_profile_gl(int m, double a) {
    if (M_PI * M_PI * (double)m * m / a > 690) return 0.0;   // the reference's underflow cut
    static double X[64], W[64];
    static bool init = false;
    if (!init) {
        gauss_legendre64(X, W);
        init = true;
    }
    double s = 0;
    for (int panel = 0; panel < 8; ++panel) {
        const double lo = -0.5 + panel / 8.0, h = 1.0 / 16;   // panel centre lo + h, half-width h
        for (int i = 0; i < 64; ++i) {
            const double k = lo + h + h * X[i];
            s += W[i] * h * std::exp(-a * k * k) * std::cos(2 * M_PI * k * m);
        }
    }
    return s;
}

void spread_kernel(double fwhm, int half, std::vector<double>& K) {
    const int n = 2 * half + 1;
    const double re = fwhm / 16 * 1.67834699;
    K.assign((size_t)n * n, 0.0);
    std::vector<double> Iv(n);
    for (int c = 0; c < 6; ++c) {
        const double a = 2 * M_PI * M_PI * re * re * EXPGAL_VAR[c];
        for (int m = -half; m <= half; ++m) Iv[m + half] = band_profile_gl(m, a);
        for (int jj = 0; jj < n; ++jj)
            for (int ii = 0; ii < n; ++ii) K[ii + jj * n] += EXPGAL_AMP[c] * Iv[ii] * Iv[jj];
    }
    double s = 0;
    for (double v : K) s += v;
    for (double& v : K) v /= s;
}

// _spread_fwhm (psf_photometry_diagnostics.jl:113-123): effective-area FWHM of
// a unit-flux render of the model (prm: full parameters at the star) over its
// integer extent (PSF.jl:285-291 with fwhm_factor 5; ImagePSF
// empirical_models.jl:167-176).  NaN when undefined.
double spread_fwhm(const Psf& P, const double* prm_star, int row_flux, int row_bkg) {
    double prm[8];
    for (int k = 0; k < 8; ++k) prm[k] = prm_star[k];
    prm[row_flux] = 1.0;
    prm[row_bkg] = 0.0;
    const double Y = prm[0], X = prm[1];
    if (!(std::isfinite(Y) && std::isfinite(X))) return NAN;
    double ylo, yhi, xlo, xhi;
    if (P.kind == CP_PSF_CIRCULAR_GAUSSIAN || P.kind == CP_PSF_CIRCULAR_MOFFAT) {
        const double fw = P.kind == CP_PSF_CIRCULAR_GAUSSIAN ? prm[2] : 2 * prm[2] * std::sqrt(std::exp2(1 / prm[3]) - 1);
        ylo = Y - 5 * fw / 2; yhi = Y + 5 * fw / 2;
        xlo = X - 5 * fw / 2; xhi = X + 5 * fw / 2;
    } else {
        ylo = Y - (P.oy - 1) / P.os_y; yhi = Y + (P.gny - P.oy) / P.os_y;
        xlo = X - (P.ox - 1) / P.os_x; xhi = X + (P.gnx - P.ox) / P.os_x;
    }
    double s1 = 0, s2 = 0;
    for (long gx = (long)std::floor(xlo); gx <= (long)std::ceil(xhi); ++gx)
        for (long gy = (long)std::floor(ylo); gy <= (long)std::ceil(yhi); ++gy) {
            const double g = evaluate(P, prm, (double)gy, (double)gx);
            s1 += g;
            s2 += g * g;
        }
    return s2 > 0 ? std::sqrt(s1 * s1 / s2 * 2 * std::log(2.0) / M_PI) : NAN;
}

// _star_diagnostics! (psf_photometry_diagnostics.jl:154-231) for one star
// (prm: its full parameters) over the box [ylo, yhi] x [xlo, xhi] (1-based,
// inside the image): data is the original image, resid the residual with this
// star's model (and its neighbours') removed, K the spread_model disk (empty:
// spread_model left NaN), p the free-parameter count.  Writes outs[k][i].
void star_diagnostics(const Psf& P, const double* prm, int row_flux, int row_bkg, int ylo, int yhi, int xlo,
                      int xhi, const double* data, const double* resid, const double* inv_var, int ny,
                      const std::vector<double>& K, int kh, int p, double* const outs[6], int i) {
    if (ylo > yhi || xlo > xhi) return;
    const int by = yhi - ylo + 1, bx = xhi - xlo + 1, kw = 2 * kh + 1;
    const bool do_spread = !K.empty();
    std::vector<double> ms((size_t)by * bx), gs((size_t)by * bx, 0.0);
    for (int ix = 0; ix < bx; ++ix)
        for (int iy = 0; iy < by; ++iy) ms[iy + ix * by] = evaluate(P, prm, ylo + iy, xlo + ix);
    if (do_spread)   // correlate!(.., :zero), correlation.jl:122-140
        for (int ix = 0; ix < bx; ++ix)
            for (int iy = 0; iy < by; ++iy) {
                double g = 0;
                for (int b = -kh; b <= kh; ++b)
                    for (int a = -kh; a <= kh; ++a) {
                        const int jy = iy + a, jx = ix + b;
                        if (jy < 0 || jy >= by || jx < 0 || jx >= bx) continue;
                        g += ms[jy + jx * by] * K[(a + kh) + (b + kh) * kw];
                    }
                gs[iy + ix * by] = g;
            }
    const double flux = prm[row_flux], bkg = prm[row_bkg];
    if (!(flux > 0)) return;
    const double inv_flux = 1 / flux;
    double qfit = 0, nclean = 0, ndirty = 0, dcrowd = 0, Sp = 0, Sq = 0, Sp0 = 0, Sq0 = 0, SqG = 0;
    double ssum = 0, s2sum = 0;
    int ngood = 0;
    for (int ix = 0; ix < bx; ++ix)
        for (int iy = 0; iy < by; ++iy) {
            const size_t q = (size_t)(ylo + iy - 1) + (size_t)(xlo + ix - 1) * ny;
            const double wp = inv_var ? inv_var[q] : 1.0;
            if (!(std::isfinite(wp) && wp > 0)) continue;
            const double rq = resid[q], mv = ms[iy + ix * by];
            qfit += std::fabs(rq);
            const double Pp = (mv - bkg) * inv_flux, wP = wp * Pp, pi = rq + mv - bkg;
            nclean += wP * pi;
            ndirty += wP * (data[q] - bkg);
            dcrowd += wP * Pp;
            if (do_spread) {
                const double gi = (gs[iy + ix * by] - bkg) * inv_flux;
                Sp += wP * pi;
                Sq += wp * gi * pi;
                Sp0 += wP * Pp;
                Sq0 += wp * Pp * gi;
                SqG += wp * gi * gi;
            }
            if (inv_var) {
                const double sg = 1 / std::sqrt(wp);
                ssum += sg;
                s2sum += sg * sg;
                ++ngood;
            }
        }
    if (outs[0]) outs[0][i] = qfit * inv_flux;
    if (outs[3] && dcrowd > 0 && nclean > 0 && ndirty > 0) outs[3][i] = 2.5 * std::log10(ndirty / nclean);
    if (do_spread && Sp > 0) {
        if (outs[4]) outs[4][i] = Sq / Sp - Sq0 / Sp0;
        if (inv_var && outs[5]) {
            const double v = Sp * Sp * SqG + Sq * Sq * Sp0 - 2 * Sq * Sp * Sq0;
            outs[5][i] = std::sqrt(std::max(0.0, v)) / (Sp * Sp);
        }
    }
    if (inv_var) {
        const double c = std::sqrt(2 / M_PI);
        if (outs[1]) outs[1][i] = c * ssum * inv_flux;
        if (outs[2] && s2sum > 0 && ngood > p) {
            const double dof_f = std::sqrt(1 - (double)p / ngood);
            outs[2][i] = (qfit - c * dof_f * ssum) / (std::sqrt((1 - 2 / M_PI) * s2sum) * dof_f);
        }
    }
}

}  // namespace

extern "C" {

// The spread_model reference disk as the oracle computes it (for tests).
int or_spread_kernel(double fwhm, int32_t half, double* K) {
    if (!(fwhm > 0) || half < 0 || half > 31 || !K) return fail("bad spread kernel arguments");
    std::vector<double> k;
    spread_kernel(fwhm, half, k);
    std::copy(k.begin(), k.end(), K);
    return 0;
}

// ---------------------------------------------------------------------------
// fit_all_stars_simultaneous (psf_photometry_simultaneous.jl:953-1431)
// ---------------------------------------------------------------------------
int or_fit_simultaneous(const cp_psf_desc* psfd, const double* image, const double* inv_var,
                        int32_t ny, int32_t nx, const double* y0, const double* x0,
                        const double* flux0, int32_t n_stars, double fit_rad,
                        const cp_sim_opts* o, cp_sim_result* out) {
    // -- argument contract (:984-997)
    if (o->max_iter <= 0) return fail("max_iter must be positive");
    if (o->inner_iterations <= 0) return fail("inner_iterations must be positive");
    if (!(o->linear_tol > 0)) return fail("linear_tol must be positive");
    if (o->max_trials <= 0) return fail("max_trials must be positive");
    double model_rad_max = o->model_rad_max > 0 ? o->model_rad_max : 10 * fit_rad;
    if (!(model_rad_max >= fit_rad)) return fail("model_rad_max must be >= fit_rad");
    if (!(o->model_rad_nsigma > 0)) return fail("model_rad_nsigma must be positive");
    if (o->solver != CP_SOLVER_LSQR && o->solver != CP_SOLVER_CGLS && o->solver != CP_SOLVER_LSMR)
        return fail("unknown solver");
    if (!(fit_rad > 0)) return fail("fit_rad must be positive");
    Psf P;
    if (make_psf(psfd, P)) return CP_EARG;
    const int np_full = nparams_kind(P.kind);
    const int row_flux = flux_slot(P.kind), row_bkg = bkg_slot(P.kind);
    // free set in getproperties order (psf_fitting.jl:8-14): y, x, flux
    std::vector<int> free_rows, grad_col;
    if (o->free_mask & CP_FREE_Y) { free_rows.push_back(0); grad_col.push_back(1); }
    if (o->free_mask & CP_FREE_X) { free_rows.push_back(1); grad_col.push_back(2); }
    if (o->free_mask & CP_FREE_FLUX) { free_rows.push_back(row_flux); grad_col.push_back(3); }
    const int p = (int)free_rows.size();
    if (p == 0) return fail("all model parameters are fixed; nothing to fit");
    int k_y = -1, k_x = -1;
    for (int k = 0; k < p; ++k) {
        if (grad_col[k] == 1) k_y = k;
        if (grad_col[k] == 2) k_x = k;
    }
    // fixed template: shape params from the descriptor, bkg and any fixed y/x/flux
    std::vector<double> tmpl(np_full, 0.0);
    if (P.kind == CP_PSF_CIRCULAR_GAUSSIAN) tmpl[2] = P.fwhm;
    if (P.kind == CP_PSF_CIRCULAR_MOFFAT) { tmpl[2] = P.alpha; tmpl[3] = P.beta; }
    tmpl[row_bkg] = o->fixed_bkg;
    const bool y_fixed = !(o->free_mask & CP_FREE_Y), x_fixed = !(o->free_mask & CP_FREE_X),
               f_fixed = !(o->free_mask & CP_FREE_FLUX);

    // catalogue -> params (psf_photometry_single.jl:132-197)
    std::vector<double> py(y0, y0 + n_stars), px(x0, x0 + n_stars), pf(flux0, flux0 + n_stars);
    std::vector<double> ey(n_stars, NAN), ex(n_stars, NAN), ef(n_stars, NAN);
    if (y_fixed) std::fill(ey.begin(), ey.end(), 0.0);
    if (x_fixed) std::fill(ex.begin(), ex.end(), 0.0);
    if (f_fixed) std::fill(ef.begin(), ef.end(), 0.0);
    const int64_t npix = (int64_t)ny * nx;
    auto write_out_common = [&](const std::vector<uint8_t>& valid) {
        for (int i = 0; i < n_stars; ++i) {
            out->y[i] = py[i]; out->x[i] = px[i]; out->flux[i] = pf[i];
            out->y_err[i] = ey[i]; out->x_err[i] = ex[i]; out->flux_err[i] = ef[i];
            out->valid[i] = valid[i];
        }
    };
    out->trace_len = 0;
    out->n_trials_total = 0;
    out->n_inner_total = 0;
    if (n_stars == 0) {
        out->n_passes = 0; out->n_failed = 0; out->n_active = 0; out->n_union = 0;
        out->cost_final = 0;
        return 0;
    }

    // -- data and weights (:1056-1084)
    std::vector<double> data(image, image + npix), w(npix);
    if (!inv_var) {
        for (int64_t q = 0; q < npix; ++q)
            if (!std::isfinite(data[q]))
                return fail("the image contains non-finite values but `inv_var` was not provided");
        std::fill(w.begin(), w.end(), 1.0);
    } else {
        for (int64_t q = 0; q < npix; ++q) {
            double wv = inv_var[q];
            w[q] = (std::isfinite(wv) && wv > 0) ? wv : 0.0;
        }
    }
    for (int64_t q = 0; q < npix; ++q)
        if (!std::isfinite(data[q])) { data[q] = 0; w[q] = 0; }

    // -- stamp footprint (:1089-1105)
    const int R = (int)std::ceil(fit_rad), S = 2 * R + 1, S2 = S * S;
    std::vector<int32_t> ancy(n_stars), ancx(n_stars), active(n_stars), pixels((size_t)S2 * n_stars);
    std::vector<uint8_t> umask(npix);
    int32_t na = 0;
    or_build_stamps(w.data(), ny, nx, y0, x0, flux0, n_stars, fit_rad, ancy.data(), ancx.data(),
                    active.data(), &na, pixels.data(), umask.data());
    std::vector<uint8_t> valid(n_stars, 0);
    for (int j = 0; j < na; ++j) valid[active[j]] = 1;
    if (out->active)
        for (int i = 0; i < n_stars; ++i) out->active[i] = valid[i];
    const int n_failed = n_stars - na;
    if (na == 0) {
        // _empty_simultaneous_result (:1433-1445)
        write_out_common(valid);
        for (int i = 0; i < n_stars; ++i) {
            out->bkg[i] = 0; out->bkg_err[i] = 0; out->converged[i] = 0; out->n_iter[i] = 0;
            if (out->chisq) out->chisq[i] = 0;
            for (double* d : {out->qfit, out->qfit_expected, out->qfit_z, out->crowding, out->spread_model,
                              out->spread_model_err})
                if (d) d[i] = NAN;
        }
        if (out->residual) std::memset(out->residual, 0, sizeof(float) * npix);
        out->n_passes = 0; out->n_failed = n_stars; out->n_active = 0; out->n_union = 0;
        out->cost_final = 0;
        return 0;
    }
    std::vector<int64_t> union_pix;
    for (int64_t q = 0; q < npix; ++q)
        if (umask[q]) union_pix.push_back(q);
    const int64_t U = (int64_t)union_pix.size();

    // -- model radii and ring table (:1113-1124)
    if (o->model_rad <= 0 && !inv_var)
        return fail("model_rad=:auto requires inv_var (it needs a noise scale)");
    int R_fit = R;
    int R_cap = (int)std::ceil(o->model_rad <= 0 ? model_rad_max : std::max(o->model_rad, fit_rad));
    std::vector<double> finit(na);
    for (int j = 0; j < na; ++j) finit[j] = flux0[active[j]];
    std::vector<int32_t> model_R(na);
    if (or_model_radii(psfd, o->model_rad, o->model_rad_nsigma, R_fit, R_cap, w.data(), npix,
                       finit.data(), na, model_R.data()))
        return CP_EARG;
    int R_max = *std::max_element(model_R.begin(), model_R.end());
    std::vector<int32_t> mdy((2 * R_max + 1) * (2 * R_max + 1)), mdx(mdy.size());
    or_ring_offsets(R_max, mdy.data(), mdx.data());
    std::vector<int> model_S2(na);
    for (int j = 0; j < na; ++j) model_S2[j] = (2 * model_R[j] + 1) * (2 * model_R[j] + 1);

    // -- working buffers (:1129-1166)
    const int64_t n = (int64_t)p * na;
    std::vector<double> theta(n);
    for (int j = 0; j < na; ++j) {
        int i = active[j];
        double full[3] = {py[i], px[i], pf[i]};
        for (int k = 0; k < p; ++k) theta[(size_t)j * p + k] = full[grad_col[k] - 1];
    }
    std::vector<double> V((size_t)p * S2 * na, 0.0), colnorm((size_t)p * na, 0.0);
    std::vector<double> model_img(npix), model_cand(npix), wt_resid(npix, 0.0), mrhs(npix);
    std::vector<double> b_scaled(n), dsc(n), dlt(n), D(n);
    std::vector<uint8_t> live(na, 1), star_conv(na, 0);
    std::vector<int> frozen_at(na, 0);
    std::vector<double> data_work = data;

    auto star_params = [&](const double* th, int j, double* prm) {
        for (int k = 0; k < np_full; ++k) prm[k] = tmpl[k];
        prm[0] = y_fixed ? o->fixed_y : 0;
        prm[1] = x_fixed ? o->fixed_x : 0;
        prm[row_flux] = f_fixed ? o->fixed_flux : 0;
        for (int k = 0; k < p; ++k) prm[free_rows[k]] = th[(size_t)j * p + k];
    };

    // _fill_stamps! (:342-399; the Gridded specialisation :437-543 computes the
    // same derivatives, so one restatement serves both)
    auto fill_stamps = [&](const std::vector<uint8_t>& lv) {
        double prm[8], G[8], f;
        for (int a = 0; a < na; ++a) {
            if (!lv[a]) continue;
            for (int k = 0; k < p; ++k) colnorm[(size_t)a * p + k] = 0;
            star_params(theta.data(), a, prm);
            int ay = ancy[active[a]], ax = ancx[active[a]];
            int mi = 0;
            for (int dx = -R; dx <= R; ++dx)
                for (int dy = -R; dy <= R; ++dy, ++mi) {
                    int32_t fi = pixels[(size_t)a * S2 + mi];
                    if (fi == 0) continue;
                    evaluate_fg(P, prm, (double)(ay + dy), (double)(ax + dx), f, G);
                    double g3[3] = {G[0], G[1], G[row_flux]};
                    double sw = std::sqrt(w[fi - 1]);
                    for (int k = 0; k < p; ++k) {
                        double raw = sw * g3[grad_col[k] - 1];
                        V[((size_t)a * S2 + mi) * p + k] = raw;
                        colnorm[(size_t)a * p + k] += raw * raw;
                    }
                }
        }
        for (int a = 0; a < na; ++a) {
            if (!lv[a]) continue;
            for (int k = 0; k < p; ++k) {
                double cn = std::sqrt(colnorm[(size_t)a * p + k]);
                cn = cn < EPS64 ? EPS64 : cn;
                colnorm[(size_t)a * p + k] = cn;
                double inv = 1.0 / cn;
                for (int mi = 0; mi < S2; ++mi) V[((size_t)a * S2 + mi) * p + k] *= inv;
            }
        }
    };

    // _render_model! (:545-577 / :605-673): scatter-add in active order over the
    // per-star model box.
    auto render = [&](std::vector<double>& img, const double* th, const std::vector<uint8_t>& lv) {
        std::fill(img.begin(), img.end(), 0.0);
        double prm[8];
        for (int a = 0; a < na; ++a) {
            if (!lv[a]) continue;
            star_params(th, a, prm);
            int ay = ancy[active[a]], ax = ancx[active[a]];
            for (int mi = 0; mi < model_S2[a]; ++mi) {
                int gy = ay + mdy[mi], gx = ax + mdx[mi];
                if (gy < 1 || gy > ny || gx < 1 || gx > nx) continue;
                img[(size_t)(gy - 1) + (size_t)(gx - 1) * ny] += evaluate(P, prm, gy, gx);
            }
        }
    };
    auto residual_cost = [&]() {  // :675-684
        double cost = 0;
        for (int64_t q : union_pix) {
            double r = model_img[q] - data_work[q];
            double wr = std::sqrt(w[q]) * r;
            wt_resid[q] = wr;
            cost += wr * wr;
        }
        return cost;
    };
    auto cost_of = [&](const std::vector<double>& img) {  // :686-693
        double cost = 0;
        for (int64_t q : union_pix) {
            double r = img[q] - data_work[q];
            cost += w[q] * r * r;
        }
        return cost;
    };
    auto freeze_star = [&](int j) {  // :739-754
        double prm[8];
        star_params(theta.data(), j, prm);
        int ay = ancy[active[j]], ax = ancx[active[j]];
        for (int mi = 0; mi < model_S2[j]; ++mi) {
            int gy = ay + mdy[mi], gx = ax + mdx[mi];
            if (gy < 1 || gy > ny || gx < 1 || gx > nx) continue;
            data_work[(size_t)(gy - 1) + (size_t)(gx - 1) * ny] -= evaluate(P, prm, gy, gx);
        }
    };

    Op op;
    op.m = npix;
    op.n = n;
    op.J = [&](const double* v, double* res) {
        apply_J_impl(res, npix, V.data(), pixels.data(), p, S2, na, v, live.data());
    };
    op.JT = [&](const double* u, double* res) {
        apply_JT_impl(res, V.data(), pixels.data(), p, S2, na, u, live.data());
    };

    // -- initial fill and cost (:1171-1176)
    fill_stamps(live);
    render(model_img, theta.data(), live);
    double cost = residual_cost();
    for (int64_t i = 0; i < n; ++i) D[i] = colnorm[i];
    const double dof = (double)std::max<int64_t>(U - n, 1);
    double lam = o->lambda_init;
    int n_run = 0;
    std::vector<double> theta_c(n);

    // -- outer damped Gauss-Newton loop (:1186-1284)
    for (int iter = 1; iter <= o->max_iter; ++iter) {
        fill_stamps(live);
        render(model_img, theta.data(), live);
        cost = residual_cost();
        for (int64_t q = 0; q < npix; ++q) mrhs[q] = -wt_resid[q];
        op.JT(wt_resid.data(), b_scaled.data());
        double C_r = cost / dof;
        double gnorm = 0;
        for (int64_t i = 0; i < n; ++i) {
            double bt = colnorm[i] * b_scaled[i];
            double Aii = colnorm[i] * colnorm[i];
            gnorm = std::max(gnorm, std::fabs(bt) / (std::sqrt(Aii * C_r) + EPS64));
        }
        if (gnorm <= o->g_tol) {
            for (int j = 0; j < na; ++j)
                if (live[j]) { live[j] = 0; frozen_at[j] = iter - 1; star_conv[j] = 1; }
            n_run = iter - 1;
            break;
        }
        bool accepted = false, f_conv = false;
        for (int trial = 1; trial <= o->max_trials; ++trial) {
            double lam_solve = std::sqrt(lam);
            const auto solve = o->solver == CP_SOLVER_CGLS ? cgls : (o->solver == CP_SOLVER_LSMR ? lsmr : lsqr);
            int its = solve(op, mrhs.data(), lam_solve, o->inner_iterations, o->linear_tol, o->linear_tol, dsc.data());
            out->n_inner_total += its;
            out->n_trials_total += 1;
            for (int64_t i = 0; i < n; ++i) dlt[i] = dsc[i] / colnorm[i];
            // _cap_position_step! (:700-716)
            if (k_y >= 0 || k_x >= 0)
                for (int a = 0; a < na; ++a) {
                    double dy = k_y >= 0 ? dlt[(size_t)a * p + k_y] : 0.0;
                    double dx = k_x >= 0 ? dlt[(size_t)a * p + k_x] : 0.0;
                    double len = std::sqrt(dy * dy + dx * dx);
                    if (!(len > o->max_step)) continue;
                    double sc = o->max_step / len;
                    if (k_y >= 0) dlt[(size_t)a * p + k_y] *= sc;
                    if (k_x >= 0) dlt[(size_t)a * p + k_x] *= sc;
                }
            for (int64_t i = 0; i < n; ++i) {
                dsc[i] = colnorm[i] * dlt[i];
                theta_c[i] = theta[i] + dlt[i];
            }
            render(model_cand, theta_c.data(), live);
            double cost_c = cost_of(model_cand);
            double dcost = cost - cost_c;
            f_conv = (cost_c < cost) && (dcost / cost <= o->f_tol);
            if (out->trace && out->trace_len < out->trace_cap) {
                int nl = 0;
                for (int j = 0; j < na; ++j) nl += live[j];
                double* t = out->trace + 8 * (size_t)out->trace_len++;
                t[0] = iter; t[1] = trial; t[2] = cost; t[3] = cost_c; t[4] = lam; t[5] = nl; t[6] = its; t[7] = gnorm;
            }
            if (cost_c < cost) {
                accepted = true;
                theta = theta_c;
                cost = cost_c;
                lam = std::max(lam / o->lambda_down, o->lambda_min);
                for (int64_t i = 0; i < n; ++i) D[i] = std::max(D[i], colnorm[i]);
                break;
            } else {
                lam = std::min(lam * o->lambda_up, o->lambda_max);
            }
        }
        n_run = iter;
        if (f_conv) {
            for (int j = 0; j < na; ++j)
                if (live[j]) { live[j] = 0; frozen_at[j] = iter; star_conv[j] = 1; }
            break;
        }
        if (accepted) {
            for (int j = 0; j < na; ++j) {
                if (!live[j]) continue;
                // _x_tol_local (:810-823)
                double num = 0, xn = 0;
                for (int k = 0; k < p; ++k) {
                    size_t i = (size_t)j * p + k;
                    double v = (D[i] / colnorm[i]) * dsc[i];
                    num += v * v;
                    double xv = D[i] * theta[i];
                    xn += xv * xv;
                }
                bool xt = std::sqrt(num) <= o->x_tol * (std::sqrt(xn) + o->x_tol);
                // _g_tol_local (:789-801)
                double gl = 0;
                for (int k = 0; k < p; ++k) {
                    size_t i = (size_t)j * p + k;
                    double bt = colnorm[i] * b_scaled[i];
                    double Aii = colnorm[i] * colnorm[i];
                    gl = std::max(gl, std::fabs(bt) / (std::sqrt(Aii * C_r) + EPS64));
                }
                bool gt = gl <= o->g_tol;
                if (xt || gt) {
                    freeze_star(j);
                    live[j] = 0;
                    frozen_at[j] = iter;
                    star_conv[j] = 1;
                }
            }
            bool any = false;
            for (int j = 0; j < na; ++j) any |= live[j] != 0;
            if (!any) break;
        }
    }

    // -- post-fit diagnostics (:1344-1397; _star_diagnostics!,
    // psf_photometry_diagnostics.jl:154-231; _spread_fwhm :113-123)
    auto diagnostics = [&](const std::vector<double>& Y, const std::vector<double>& X,
                           const std::vector<double>& Fl, const std::vector<uint8_t>& vld) {
        double* outs[6] = {out->qfit, out->qfit_expected, out->qfit_z, out->crowding, out->spread_model,
                           out->spread_model_err};
        for (double* d : outs)
            if (d)
                for (int i = 0; i < n_stars; ++i) d[i] = NAN;
        bool any = false;
        for (double* d : outs) any = any || d;
        if (!any) return;
        // reference disk: explicit FWHM, else the effective-area FWHM of a
        // unit-flux render of the model at catalogue star 1's fitted position
        double prm1[8];
        for (int k = 0; k < np_full; ++k) prm1[k] = tmpl[k];
        prm1[0] = Y[0];
        prm1[1] = X[0];
        const double sfw = o->spread_model_fwhm > 0 ? o->spread_model_fwhm : spread_fwhm(P, prm1, row_flux, row_bkg);
        const int kh = (int)std::ceil(fit_rad);
        std::vector<double> K;
        if (std::isfinite(sfw) && sfw > 0) spread_kernel(sfw, kh, K);
        std::vector<double> resid((size_t)ny * nx);
        for (size_t q = 0; q < resid.size(); ++q) resid[q] = data[q] - model_img[q];
        double prm[8];
        for (int j = 0; j < na; ++j) {
            const int i = active[j];
            if (!vld[i]) continue;
            for (int k = 0; k < np_full; ++k) prm[k] = tmpl[k];
            prm[0] = Y[i];
            prm[1] = X[i];
            prm[row_flux] = Fl[i];
            // +-fit_rad box at the fitted position, clamped (:1361-1365)
            const int ylo = std::max((int)std::floor(Y[i] - fit_rad), 1);
            const int yhi = std::min((int)std::ceil(Y[i] + fit_rad), (int)ny);
            const int xlo = std::max((int)std::floor(X[i] - fit_rad), 1);
            const int xhi = std::min((int)std::ceil(X[i] + fit_rad), (int)nx);
            star_diagnostics(P, prm, row_flux, row_bkg, ylo, yhi, xlo, xhi, data.data(), resid.data(), inv_var, ny, K,
                             kh, p, outs, i);
        }
    };

    // -- final bookkeeping and errors (:1289-1329)
    for (int j = 0; j < na; ++j) {
        int i = active[j];
        for (int k = 0; k < p; ++k) {
            double v = theta[(size_t)j * p + k];
            if (grad_col[k] == 1) py[i] = v;
            if (grad_col[k] == 2) px[i] = v;
            if (grad_col[k] == 3) pf[i] = v;
        }
    }
    if (y_fixed) std::fill(py.begin(), py.end(), o->fixed_y);
    if (x_fixed) std::fill(px.begin(), px.end(), o->fixed_x);
    if (f_fixed) std::fill(pf.begin(), pf.end(), o->fixed_flux);
    std::vector<uint8_t> all(na, 1);
    fill_stamps(all);
    int ce = o->covariance;
    if (ce == CP_COV_DEFAULT) ce = inv_var ? CP_COV_KNOWN : CP_COV_REWEIGHTED;
    for (int j = 0; j < na; ++j) {
        int i = active[j];
        if (ce == CP_COV_NONE) continue;
        std::vector<double> blk(p * p, 0.0), cov(p * p);
        for (int mi = 0; mi < S2; ++mi) {
            if (pixels[(size_t)j * S2 + mi] == 0) continue;
            for (int k = 0; k < p; ++k)
                for (int l = 0; l < p; ++l)
                    blk[k * p + l] += V[((size_t)j * S2 + mi) * p + k] * V[((size_t)j * S2 + mi) * p + l];
        }
        for (int k = 0; k < p; ++k)
            for (int l = 0; l < p; ++l) blk[k * p + l] *= colnorm[(size_t)j * p + k] * colnorm[(size_t)j * p + l];
        double tr = 0;
        for (int k = 0; k < p; ++k) tr += blk[k * p + k];
        for (int k = 0; k < p; ++k) blk[k * p + k] += 1e-12 * tr;
        sym_inverse(blk.data(), p, cov.data());
        if (ce == CP_COV_REWEIGHTED)
            for (auto& c : cov) c *= cost / dof;
        for (int k = 0; k < p; ++k) {
            double e = std::sqrt(std::max(0.0, cov[k * p + k]));
            if (grad_col[k] == 1) ey[i] = e;
            if (grad_col[k] == 2) ex[i] = e;
            if (grad_col[k] == 3) ef[i] = e;
        }
    }
    // validity gate (:1332-1337)
    for (int i = 0; i < n_stars; ++i)
        if (valid[i] && !(std::isfinite(py[i]) && std::isfinite(px[i]) && std::isfinite(pf[i]) && pf[i] > 0))
            valid[i] = 0;
    // final render against the original data, residual, chisq (:1344-1397)
    render(model_img, theta.data(), all);
    if (out->residual)
        for (int64_t q = 0; q < npix; ++q) out->residual[q] = (float)(data[q] - model_img[q]);
    write_out_common(valid);
    for (int i = 0; i < n_stars; ++i) {
        out->bkg[i] = o->fixed_bkg;
        out->bkg_err[i] = 0;
        out->converged[i] = 0;
        out->n_iter[i] = 0;
        if (out->chisq) out->chisq[i] = 0;
    }
    for (int j = 0; j < na; ++j) {
        int i = active[j];
        out->converged[i] = star_conv[j];
        out->n_iter[i] = live[j] ? n_run : frozen_at[j];
        if (!out->chisq || !valid[i]) continue;
        int ylo = (int)std::floor(py[i] - fit_rad), yhi = (int)std::ceil(py[i] + fit_rad);
        int xlo = (int)std::floor(px[i] - fit_rad), xhi = (int)std::ceil(px[i] + fit_rad);
        ylo = std::max(ylo, 1); yhi = std::min(yhi, (int)ny);
        xlo = std::max(xlo, 1); xhi = std::min(xhi, (int)nx);
        if (ylo > yhi || xlo > xhi) continue;
        double num = 0;
        int good = 0;
        for (int gx = xlo; gx <= xhi; ++gx)
            for (int gy = ylo; gy <= yhi; ++gy) {
                size_t q = (size_t)(gy - 1) + (size_t)(gx - 1) * ny;
                double iv = inv_var ? inv_var[q] : 1.0;
                if (std::isfinite(iv) && iv > 0) {
                    double r = data[q] - model_img[q];
                    num += iv * r * r;
                    ++good;
                }
            }
        int den = good - p;
        if (den > 0) out->chisq[i] = num / den;
    }
    diagnostics(py, px, pf, valid);
    out->n_passes = n_run;
    out->n_failed = n_failed;
    out->n_active = na;
    out->n_union = U;
    out->cost_final = cost;
    return 0;
}

// _extract_errors! (psf_photometry_single.jl:210-218); free_idx 0-based rows.
int or_extract_errors(const double* cov, int32_t p, const int32_t* free_idx, int32_t n_params,
                      double* errors_col) {
    std::vector<char> is_free(n_params, 0);
    for (int j = 0; j < p; ++j) {
        errors_col[free_idx[j]] = std::sqrt(std::max(0.0, cov[j * p + j]));
        is_free[free_idx[j]] = 1;
    }
    for (int k = 0; k < n_params; ++k)
        if (!is_free[k]) errors_col[k] = 0;
    return 0;
}

// The line-fit problem of test/psf/psf_fitting_tests.jl:143-175.
int or_lm_line(const double* xs, const double* ys, int32_t n, const double* x0,
               const cp_lm_opts* o, double* x_out, cp_lm_stat* st) {
    std::vector<double> x(x0, x0 + 2);
    Accum acc = [&](double* A, double* b, const double* xv) {
        std::fill(A, A + 4, 0.0);
        std::fill(b, b + 2, 0.0);
        double cost = 0;
        for (int k = 0; k < n; ++k) {
            double r = xv[0] * xs[k] + xv[1] - ys[k];
            cost = std::fma(r, r, cost);
            A[0] = std::fma(xs[k], xs[k], A[0]);
            A[2] = std::fma(1.0, xs[k], A[2]);
            A[1] = std::fma(xs[k], 1.0, A[1]);
            A[3] += 1;
            b[0] = std::fma(r, xs[k], b[0]);
            b[1] += r;
        }
        return cost;
    };
    int rc = lm_irls(acc, 2, n, false, x, *o, *st, nullptr);
    x_out[0] = x[0];
    x_out[1] = x[1];
    return rc;
}

namespace {
// fit_star over a column-major ny x nx block whose pixel (i, j) sits at image
// coordinates (oy + i, ox + j) (psf_fitting.jl:229-308 with
// _prepare_fit_star_inputs :36-101); generic accum! over evaluate_fg.
int fit_star_block(const cp_psf_desc* psfd, const double* pix, int ny, int nx, int oy, int ox,
                   const double* inv_var, const double* theta0, uint32_t free_mask, const cp_lm_opts* o,
                   double* theta_out, double* cov_out, cp_lm_stat* st) {
    Psf P;
    if (make_psf(psfd, P)) return CP_EARG;
    const int npf = nparams_kind(P.kind);
    std::vector<int> free_idx;
    for (int k = 0; k < npf; ++k)
        if (free_mask & (1u << k)) free_idx.push_back(k);
    const int p = (int)free_idx.size();
    if (p == 0) return fail("all model parameters are fixed; nothing to fit");
    const int nobs = ny * nx;
    std::vector<double> wts;
    if (inv_var) {
        int n_valid = 0;
        wts.resize(nobs);
        for (int q = 0; q < nobs; ++q) {
            double wv = inv_var[q];
            bool ok = std::isfinite(wv) && wv > 0;
            wts[q] = ok ? wv : 0.0;
            n_valid += ok;
        }
        if (!(n_valid > p)) return fail("too few valid inv_var pixels");
    }
    std::vector<double> full(theta0, theta0 + npf);
    std::vector<double> x(p);
    for (int k = 0; k < p; ++k) x[k] = full[free_idx[k]];
    Accum acc = [&](double* A, double* b, const double* xv) {
        std::fill(A, A + p * p, 0.0);
        std::fill(b, b + p, 0.0);
        std::vector<double> prm = full;
        for (int k = 0; k < p; ++k) prm[free_idx[k]] = xv[k];
        double cost = 0, f, G[8];
        int q = 0;
        for (int j = 1; j <= nx; ++j)
            for (int i = 1; i <= ny; ++i, ++q) {
                double wv = inv_var ? wts[q] : 1.0;
                evaluate_fg(P, prm.data(), (double)(oy + i), (double)(ox + j), f, G);
                double r = f - pix[q];
                double wr = wv * r;
                cost = std::fma(wr, r, cost);
                for (int jj = 0; jj < p; ++jj) {
                    double gj = G[free_idx[jj]];
                    b[jj] = std::fma(wr, gj, b[jj]);
                    for (int ii = 0; ii < p; ++ii) A[ii * p + jj] = std::fma(wv * G[free_idx[ii]], gj, A[ii * p + jj]);
                }
            }
        return cost;
    };
    int rc = lm_irls(acc, p, nobs, inv_var != nullptr, x, *o, *st, cov_out);
    if (rc) return rc;
    for (int k = 0; k < npf; ++k) theta_out[k] = full[k];
    for (int k = 0; k < p; ++k) theta_out[free_idx[k]] = x[k];
    return 0;
}
}  // namespace

int or_fit_star(const cp_psf_desc* psfd, const double* pix, int32_t ny, int32_t nx, const double* inv_var,
                const double* theta0, uint32_t free_mask, const cp_lm_opts* o, double* theta_out, double* cov_out,
                cp_lm_stat* st) {
    return fit_star_block(psfd, pix, ny, nx, 0, 0, inv_var, theta0, free_mask, o, theta_out, cov_out, st);
}

// fit_all_stars (psf_photometry_single.jl:267-469): stars sorted by flux
// (stable, brightest first) are fitted one at a time with fit_star on the
// progressively subtracted residual over their +-fit_rad box; passes after
// the first add each star's current model back before refitting it, and the
// last pass computes the per-star diagnostics right after the star's
// subtraction.  theta0: n x npf initial parameters (getproperties order),
// free_mask: free parameters; a fixed parameter k fits at fixed_values[k]
// (model_from_vector with `fixed`), or at the star's own value when
// fixed_values is NULL or holds NaN there.  Fills y, x, flux, bkg, their
// errors, converged, valid, chisq, n_iter, the diagnostics and the residual.
int or_fit_all_stars(const cp_psf_desc* psfd, const double* image, const double* inv_var, int32_t ny, int32_t nx,
                     const double* theta0, int32_t n, double fit_rad, int32_t n_passes, uint32_t free_mask,
                     const double* fixed_values, const cp_lm_opts* o, double spread_model_fwhm,
                     cp_sim_result* out) {
    Psf P;
    if (make_psf(psfd, P)) return CP_EARG;
    if (n_passes <= 0) return fail("n_passes must be positive");
    if (!(fit_rad > 0)) return fail("fit_rad must be positive");
    const int npf = nparams_kind(P.kind), rf = flux_slot(P.kind), rb = bkg_slot(P.kind);
    std::vector<int> free_idx;
    for (int k = 0; k < npf; ++k)
        if (free_mask & (1u << k)) free_idx.push_back(k);
    const int p = (int)free_idx.size();
    if (p == 0) return fail("all model parameters are fixed; nothing to fit");
    if (n == 0) {   // the empty result (:285-288): no passes, 0 x 0 residual
        out->n_passes = 0;
        out->n_failed = 0;
        return 0;
    }
    std::vector<double> prm(theta0, theta0 + (size_t)n * npf), err((size_t)n * npf, NAN);
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < npf; ++k)
            if (!(free_mask & (1u << k))) err[(size_t)i * npf + k] = 0.0;
    std::vector<double> R(image, image + (size_t)ny * nx);
    std::vector<uint8_t> valid(n, 1), conv(n, 0);
    std::vector<double> chisq(n, 0.0);
    std::vector<int> niter(n, 0);
    double* outs[6] = {out->qfit, out->qfit_expected, out->qfit_z, out->crowding, out->spread_model,
                       out->spread_model_err};
    for (double* d : outs)
        if (d)
            for (int i = 0; i < n; ++i) d[i] = NAN;
    const int kh = (int)std::ceil(fit_rad);
    std::vector<double> K;
    if (n > 0) {
        const double sfw = spread_model_fwhm > 0 ? spread_model_fwhm : spread_fwhm(P, &prm[0], rf, rb);
        if (std::isfinite(sfw) && sfw > 0) spread_kernel(sfw, kh, K);
    }
    int n_failed = 0;
    std::vector<int> order(n);
    std::vector<double> cut, ivc, th(npf), th0(npf), cov((size_t)p * p);
    auto render = [&](const double* m, int ylo, int yhi, int xlo, int xhi, double sign) {
        for (int gx = xlo; gx <= xhi; ++gx)
            for (int gy = ylo; gy <= yhi; ++gy)
                R[(size_t)(gy - 1) + (size_t)(gx - 1) * ny] += sign * evaluate(P, m, gy, gx);
    };
    for (int pass = 1; pass <= n_passes; ++pass) {
        for (int i = 0; i < n; ++i) order[i] = i;
        auto key = [&](int i) {   // sortperm(rev=true) puts NaN first
            const double f = prm[(size_t)i * npf + rf];
            return std::isnan(f) ? INFINITY : f;
        };
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return key(a) > key(b); });
        for (int idx : order) {
            if (!valid[idx]) continue;
            double* m = &prm[(size_t)idx * npf];
            // floor(Int, .) of a non-finite or huge position throws outside the try
            if (!(std::fabs(m[0]) < 1e9 && std::fabs(m[1]) < 1e9)) return fail("non-finite star position");
            const int ylo = std::max((int)std::floor(m[0] - fit_rad), 1), yhi = std::min((int)std::ceil(m[0] + fit_rad), (int)ny);
            const int xlo = std::max((int)std::floor(m[1] - fit_rad), 1), xhi = std::min((int)std::ceil(m[1] + fit_rad), (int)nx);
            const long by = (long)yhi - ylo + 1, bx = (long)xhi - xlo + 1;
            if (by <= 0 || bx <= 0 || by * bx < 3) {
                valid[idx] = 0;
                continue;
            }
            if (pass > 1) render(m, ylo, yhi, xlo, xhi, 1.0);
            cut.assign((size_t)by * bx, 0.0);
            if (inv_var) ivc.assign((size_t)by * bx, 0.0);
            for (int ix = 0; ix < bx; ++ix)
                for (int iy = 0; iy < by; ++iy) {
                    const size_t q = (size_t)(ylo + iy - 1) + (size_t)(xlo + ix - 1) * ny;
                    cut[iy + ix * by] = R[q];
                    if (inv_var) ivc[iy + ix * by] = inv_var[q];
                }
            for (int k = 0; k < npf; ++k)
                th0[k] = (!(free_mask & (1u << k)) && fixed_values && !std::isnan(fixed_values[k])) ? fixed_values[k]
                                                                                                    : m[k];
            cp_lm_stat st{};
            const int rc = fit_star_block(psfd, cut.data(), (int)by, (int)bx, ylo - 1, xlo - 1,
                                          inv_var ? ivc.data() : nullptr, th0.data(), free_mask, o, th.data(),
                                          cov.data(), &st);
            if (rc != 0) {   // the reference's catch: undo the add-back, mark invalid
                if (pass > 1) render(m, ylo, yhi, xlo, xhi, -1.0);
                valid[idx] = 0;
                ++n_failed;
                continue;
            }
            for (int k = 0; k < npf; ++k) m[k] = th[k];
            for (int j = 0; j < p; ++j) err[(size_t)idx * npf + free_idx[j]] = std::sqrt(std::max(0.0, cov[(size_t)j * p + j]));
            conv[idx] = st.converged;
            chisq[idx] = st.chisq;
            niter[idx] += st.iterations;
            render(m, ylo, yhi, xlo, xhi, -1.0);
            if (pass == n_passes)
                star_diagnostics(P, m, rf, rb, ylo, yhi, xlo, xhi, image, R.data(), inv_var, ny, K, kh, p, outs, idx);
        }
        for (int i = 0; i < n; ++i) {   // between-pass validation
            const double* m = &prm[(size_t)i * npf];
            valid[i] = valid[i] && conv[i] && std::isfinite(m[0]) && std::isfinite(m[1]) && std::isfinite(m[rf]) &&
                       m[rf] > 0;
        }
    }
    for (int i = 0; i < n; ++i) {
        const double* m = &prm[(size_t)i * npf];
        const double* e = &err[(size_t)i * npf];
        out->y[i] = m[0];
        out->x[i] = m[1];
        out->flux[i] = m[rf];
        out->bkg[i] = m[rb];
        out->y_err[i] = e[0];
        out->x_err[i] = e[1];
        out->flux_err[i] = e[rf];
        out->bkg_err[i] = e[rb];
        out->converged[i] = conv[i];
        out->valid[i] = valid[i];
        if (out->chisq) out->chisq[i] = chisq[i];
        if (out->n_iter) out->n_iter[i] = niter[i];
    }
    if (out->residual)
        for (size_t q = 0; q < R.size(); ++q) out->residual[q] = (float)R[q];
    out->n_passes = n_passes;
    out->n_failed = n_failed;
    return 0;
}

int or_fit_star_batch(const cp_psf_desc* psf, const float* pix, const int64_t* offsets,
                      const int32_t* ny, const int32_t* nx, const float* inv_var,
                      const double* theta0, uint32_t free_mask, int32_t n_cutouts,
                      const cp_lm_opts* o, double* theta_out, double* cov_out, cp_lm_stat* st,
                      int32_t nthreads) {
    Psf P;
    if (make_psf(psf, P)) return CP_EARG;
    const int npf = nparams_kind(P.kind);
    int p = 0;
    for (int k = 0; k < npf; ++k) p += (free_mask >> k) & 1u;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
#endif
    for (int c = 0; c < n_cutouts; ++c) {
        int64_t npx = (int64_t)ny[c] * nx[c];
        std::vector<double> img(npx), iv;
        for (int64_t q = 0; q < npx; ++q) img[q] = pix[offsets[c] + q];
        if (inv_var) {
            iv.resize(npx);
            for (int64_t q = 0; q < npx; ++q) iv[q] = inv_var[offsets[c] + q];
        }
        int rc = or_fit_star(psf, img.data(), ny[c], nx[c], inv_var ? iv.data() : nullptr,
                             theta0 + (size_t)c * npf, free_mask, o, theta_out + (size_t)c * npf,
                             cov_out ? cov_out + (size_t)c * p * p : nullptr, st + c);
        if (rc) {
            st[c].status = rc;
            for (int k = 0; k < npf; ++k) theta_out[(size_t)c * npf + k] = theta0[(size_t)c * npf + k];
        }
    }
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Detection: correlate (src/correlation.jl), findlocalmaxima and
// matched_filter (src/detection.jl)
// ---------------------------------------------------------------------------
namespace {

// out[row, col] = sum_{j, i} img[row + i - r - 1, col + j - c - 1] K[i, j]
// (the direct form of correlation.jl:329-346 that its tests take as ground
// truth), out-of-image pixels replicated or zero.  Kernel-order sums.
void correlate_direct(const double* img, int ny, int nx, const double* K, int kr, int kc, int border, double* out) {
    const int r = kr / 2, c = kc / 2;
    for (int col = 1; col <= nx; ++col)
        for (int row = 1; row <= ny; ++row) {
            double acc = 0;
            for (int j = 1; j <= kc; ++j)
                for (int i = 1; i <= kr; ++i) {
                    int ir = row + i - r - 1, ic = col + j - c - 1;
                    double v;
                    if (border == CP_BORDER_REPLICATE) {
                        ir = std::min(std::max(ir, 1), ny);
                        ic = std::min(std::max(ic, 1), nx);
                        v = img[(size_t)(ir - 1) + (size_t)(ic - 1) * ny];
                    } else {
                        v = (ir >= 1 && ir <= ny && ic >= 1 && ic <= nx) ? img[(size_t)(ir - 1) + (size_t)(ic - 1) * ny]
                                                                        : 0.0;
                    }
                    acc += v * K[(size_t)(i - 1) + (size_t)(j - 1) * kr];
                }
            out[(size_t)(row - 1) + (size_t)(col - 1) * ny] = acc;
        }
}

// findlocalmaxima (detection.jl:240-324): interior pixels column by column,
// then (edges) the top / bottom rows, the left / right columns and the four
// corners, each against its in-image neighbours, strictly.  -1 for the
// reference's BoundsError (a 1-pixel-wide image of length >= 3 with edges).
int local_maxima(const double* a, int H, int W, bool edges, std::vector<std::pair<int, int>>& out) {
    out.clear();
    if (H < 3 && W < 3) return 0;
    auto v = [&](int r, int c) { return a[(size_t)(r - 1) + (size_t)(c - 1) * H]; };
    for (int col = 2; col <= W - 1; ++col)
        for (int row = 2; row <= H - 1; ++row) {
            const double x = v(row, col);
            if (x > v(row - 1, col - 1) && x > v(row - 1, col) && x > v(row - 1, col + 1) && x > v(row, col - 1) &&
                x > v(row, col + 1) && x > v(row + 1, col - 1) && x > v(row + 1, col) && x > v(row + 1, col + 1))
                out.push_back({row, col});
        }
    if (!edges) return 0;
    if ((H == 1 && W >= 3) || (W == 1 && H >= 3)) return -1;
    for (int col = 2; col <= W - 1; ++col) {
        double x = v(1, col);
        if (x > v(1, col - 1) && x > v(1, col + 1) && x > v(2, col - 1) && x > v(2, col) && x > v(2, col + 1))
            out.push_back({1, col});
        x = v(H, col);
        if (x > v(H - 1, col - 1) && x > v(H - 1, col) && x > v(H - 1, col + 1) && x > v(H, col - 1) &&
            x > v(H, col + 1))
            out.push_back({H, col});
    }
    for (int row = 2; row <= H - 1; ++row) {
        double x = v(row, 1);
        if (x > v(row - 1, 1) && x > v(row - 1, 2) && x > v(row, 2) && x > v(row + 1, 1) && x > v(row + 1, 2))
            out.push_back({row, 1});
        x = v(row, W);
        if (x > v(row - 1, W - 1) && x > v(row - 1, W) && x > v(row, W - 1) && x > v(row + 1, W - 1) &&
            x > v(row + 1, W))
            out.push_back({row, W});
    }
    if (H >= 2 && W >= 2) {
        if (v(1, 1) > v(1, 2) && v(1, 1) > v(2, 1) && v(1, 1) > v(2, 2)) out.push_back({1, 1});
        if (v(1, W) > v(1, W - 1) && v(1, W) > v(2, W - 1) && v(1, W) > v(2, W)) out.push_back({1, W});
        if (v(H, 1) > v(H - 1, 1) && v(H, 1) > v(H - 1, 2) && v(H, 1) > v(H, 2)) out.push_back({H, 1});
        if (v(H, W) > v(H - 1, W - 1) && v(H, W) > v(H - 1, W) && v(H, W) > v(H, W - 1)) out.push_back({H, W});
    }
    return 0;
}

}  // namespace

int or_correlate(const double* img, int32_t ny, int32_t nx, const double* kernel, int32_t kr, int32_t kc,
                 int32_t border, double* out) {
    if (!(kr % 2 == 1 && kc % 2 == 1)) return fail("kernel dimensions must be odd");
    if (border != CP_BORDER_REPLICATE && border != CP_BORDER_ZERO) return fail("unknown border mode");
    correlate_direct(img, ny, nx, kernel, kr, kc, border, out);
    return 0;
}

int or_findlocalmaxima(const double* img, int32_t ny, int32_t nx, int32_t edges, int32_t cap, int32_t* py,
                       int32_t* px, int32_t* n) {
    std::vector<std::pair<int, int>> pk;
    if (local_maxima(img, ny, nx, edges != 0, pk)) return fail("image too narrow for edge maxima");
    *n = (int32_t)pk.size();
    for (int k = 0; k < (int)pk.size() && k < cap; ++k) {
        py[k] = pk[k].first;
        px[k] = pk[k].second;
    }
    return 0;
}

// matched_filter (detection.jl:87-205): kernel normalisation (zero-sum
// (P - mean) / (sum P^2 - (sum P)^2 / N) or P / sum P^2, degenerate
// denominators keep P with norm 1), smoothed = correlate(image, K); with
// inv_var (non-finite -> 0): num = correlate(w v, K) (w = 0 -> 0),
// den = correlate(w, K^2), significance = num / sqrt(den) (0 where den <= 0);
// without: significance = smoothed * kernel_norm.  Peaks: local maxima of the
// significance away from the edges with significance >= sigma; flux
// smoothed[p], or significance / (sqrt(den) kernel_norm^2) with inv_var.
int or_matched_filter(const double* image, const double* inv_var, int32_t ny, int32_t nx, const double* kernel,
                      int32_t kr, int32_t kc, int32_t normalize_zerosum, double sigma, int32_t border,
                      cp_mf_result* out) {
    if (!(kr % 2 == 1 && kc % 2 == 1)) return fail("kernel dimensions must be odd");
    if (border != CP_BORDER_REPLICATE && border != CP_BORDER_ZERO) return fail("unknown border mode");
    const size_t P = (size_t)ny * nx, NK = (size_t)kr * kc;
    double sumP = 0, sumP2 = 0;
    for (size_t k = 0; k < NK; ++k) {
        sumP += kernel[k];
        sumP2 += kernel[k] * kernel[k];
    }
    std::vector<double> K(kernel, kernel + NK);
    double knorm = 1.0;
    if (normalize_zerosum) {
        const double denom = sumP2 - sumP * sumP / (double)NK;
        if (denom > 0) {
            for (size_t k = 0; k < NK; ++k) K[k] = (kernel[k] - sumP / (double)NK) / denom;
            knorm = std::sqrt(denom);
        }
    } else if (sumP2 > 0) {
        for (size_t k = 0; k < NK; ++k) K[k] = kernel[k] / sumP2;
        knorm = std::sqrt(sumP2);
    }
    std::vector<double> sm(P), sig(P), den;
    correlate_direct(image, ny, nx, K.data(), kr, kc, border, sm.data());
    if (inv_var) {
        std::vector<double> w(P), wv(P), K2(NK), num(P);
        for (size_t q = 0; q < P; ++q) {
            w[q] = std::isfinite(inv_var[q]) ? inv_var[q] : 0.0;
            wv[q] = w[q] == 0 ? 0.0 : w[q] * image[q];
        }
        for (size_t k = 0; k < NK; ++k) K2[k] = K[k] * K[k];
        correlate_direct(wv.data(), ny, nx, K.data(), kr, kc, border, num.data());
        den.resize(P);
        correlate_direct(w.data(), ny, nx, K2.data(), kr, kc, border, den.data());
        for (size_t q = 0; q < P; ++q) sig[q] = den[q] > 0 ? num[q] / std::sqrt(den[q]) : 0.0;
    } else {
        for (size_t q = 0; q < P; ++q) sig[q] = sm[q] * knorm;
    }
    std::vector<std::pair<int, int>> pk;
    local_maxima(sig.data(), ny, nx, false, pk);
    int np = 0;
    for (auto& p : pk) {
        const size_t q = (size_t)(p.first - 1) + (size_t)(p.second - 1) * ny;
        if (!(sig[q] >= sigma)) continue;
        if (np < out->peaks_cap) {
            out->peak_y[np] = p.first;
            out->peak_x[np] = p.second;
            out->peak_significance[np] = sig[q];
            out->peak_flux[np] = inv_var ? (den[q] > 0 ? sig[q] / (std::sqrt(den[q]) * knorm * knorm) : 0.0) : sm[q];
        }
        ++np;
    }
    out->n_peaks = np;
    out->kernel_norm = knorm;
    if (out->kernel) std::copy(K.begin(), K.end(), out->kernel);
    if (out->smoothed) std::copy(sm.begin(), sm.end(), out->smoothed);
    if (out->significance) std::copy(sig.begin(), sig.end(), out->significance);
    if (out->smoothed_inv_var && inv_var) std::copy(den.begin(), den.end(), out->smoothed_inv_var);
    return 0;
}
