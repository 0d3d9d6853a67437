// simulate.cpp -- seeded synthetic inputs with the semantics of
// simulate_image (src/simulation.jl:211-498): source positions and fluxes
// (_sample_flux :1-28), rendering (render_sources! :321-347) and noise
// (_rand_poisson :349-365, add_noise! :409-435).
//
// Random streams are xoshiro256** seeded through splitmix64; the image is cut
// into 64k-pixel blocks, each with its own stream, so noise generation runs
// in parallel and gives the same image for any thread count.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/crowdphot_hip.h"

namespace {

struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        uint64_t z = seed;
        for (int i = 0; i < 4; ++i) {
            z += 0x9e3779b97f4a7c15ULL;
            uint64_t x = z;
            x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
            x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
            s[i] = x ^ (x >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        uint64_t r = rotl(s[1] * 5, 7) * 9;
        uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (next() >> 11) * 0x1.0p-53; }  // [0, 1)
    bool has_spare = false;
    double spare = 0;
    double normal() {  // Marsaglia polar method
        if (has_spare) {
            has_spare = false;
            return spare;
        }
        double u, v, s2;
        do {
            u = 2 * uniform() - 1;
            v = 2 * uniform() - 1;
            s2 = u * u + v * v;
        } while (s2 >= 1 || s2 == 0);
        double m = std::sqrt(-2 * std::log(s2) / s2);
        spare = v * m;
        has_spare = true;
        return u * m;
    }
};

// _sample_flux (src/simulation.jl:1-28): 0 uniform, 1 loguniform, 2 powerlaw.
double sample_flux(Rng& r, int dist, double lo, double hi, double power) {
    double u = r.uniform();
    if (dist == 0) return lo + u * (hi - lo);
    if (dist == 1 || std::fabs(power - 1) < 1e-15) return std::exp(std::log(lo) + u * (std::log(hi) - std::log(lo)));
    double q = 1 - power;
    return std::pow(std::pow(lo, q) + u * (std::pow(hi, q) - std::pow(lo, q)), 1 / q);
}

// _rand_poisson (src/simulation.jl:349-365): Knuth below 30, rounded normal above.
int64_t rand_poisson(Rng& r, double lam) {
    lam = std::max(lam, 0.0);
    if (lam < 30) {
        double L = std::exp(-lam), p = 1.0;
        int64_t k = 0;
        while (p > L) {
            ++k;
            p *= r.uniform();
        }
        return k - 1;
    }
    return std::max<int64_t>(0, (int64_t)std::nearbyint(lam + std::sqrt(lam) * r.normal()));
}

// ---- host model evaluation (float64), same formulas as the device models
const double GAUSS_PRE = -4.0 * std::log(2.0);

struct HostModel {
    int kind;
    double fwhm, alpha, beta;
    int n_nodes, gny, gnx;
    double osx, osy, ox, oy, fill;
    const float* grids;
    std::vector<double> yg, xg;
    std::vector<int> ig;
};

void cubic4v(double f1, double f2, double f3, double f4, double t, double& v) {
    double c1 = (f3 - f1) / 2, c4 = f3 - f2 - c1, c2 = 3 * c4 - (f4 - f2) / 2 + c1, c3 = c4 - c2;
    v = t * (t * (t * c3 + c2) + c1) + f2;
}

double bicubic(const float* g, int ny, int nx, double y, double x, double fill) {
    bool inb = std::isfinite(x) && std::isfinite(y) && x >= 1 && x <= nx && y >= 1 && y <= ny;
    if (!inb) return fill;
    int lx = std::min(std::max((int)std::floor(x), 1), nx - 1), ly = std::min(std::max((int)std::floor(y), 1), ny - 1);
    double dx = x - lx, dy = y - ly;
    int ix[4] = {std::max(lx - 1, 1), lx, lx + 1, std::min(lx + 2, nx)};
    int iy[4] = {std::max(ly - 1, 1), ly, ly + 1, std::min(ly + 2, ny)};
    double row[4];
    for (int r = 0; r < 4; ++r) {
        const float* c = g + (iy[r] - 1);
        cubic4v(c[(size_t)(ix[0] - 1) * ny], c[(size_t)(ix[1] - 1) * ny], c[(size_t)(ix[2] - 1) * ny],
                c[(size_t)(ix[3] - 1) * ny], dx, row[r]);
    }
    double v;
    cubic4v(row[0], row[1], row[2], row[3], dy, v);
    return v;
}

double eval_model(const HostModel& M, double Y, double X, double F, double py, double px) {
    if (M.kind == CP_PSF_CIRCULAR_GAUSSIAN) {
        double dx = px - X, dy = py - Y;
        double sq = (dx * dx + dy * dy) / (M.fwhm * M.fwhm);
        double norm = -(M_PI * M.fwhm * M.fwhm / GAUSS_PRE);
        return (F / norm) * std::exp(GAUSS_PRE * sq);
    }
    if (M.kind == CP_PSF_CIRCULAR_MOFFAT) {
        double r2 = (px - X) * (px - X) + (py - Y) * (py - Y);
        double amp = F * (M.beta - 1) / (M_PI * M.alpha * M.alpha);
        return amp * std::pow(1 + r2 / (M.alpha * M.alpha), -M.beta);
    }
    double u = M.osx * (px - X) + M.ox, v = M.osy * (py - Y) + M.oy;
    const size_t nn = (size_t)M.gny * M.gnx;
    if (M.n_nodes == 1) return F * bicubic(M.grids, M.gny, M.gnx, v, u, M.fill);
    // bilinear node blend (gridded_psf.jl:176-250)
    int ny = (int)M.yg.size(), nx = (int)M.xg.size();
    int iy = (int)(std::upper_bound(M.yg.begin(), M.yg.end(), Y) - M.yg.begin());
    int ix = (int)(std::upper_bound(M.xg.begin(), M.xg.end(), X) - M.xg.begin());
    iy = std::min(std::max(iy, 1), ny - 1);
    ix = std::min(std::max(ix, 1), nx - 1);
    double y0 = M.yg[iy - 1], y1 = M.yg[iy], x0 = M.xg[ix - 1], x1 = M.xg[ix];
    double Yc = std::min(std::max(Y, y0), y1), Xc = std::min(std::max(X, x0), x1);
    double inv = 1 / ((x1 - x0) * (y1 - y0));
    double w[4] = {(x1 - Xc) * (y1 - Yc) * inv, (Xc - x0) * (y1 - Yc) * inv, (x1 - Xc) * (Yc - y0) * inv,
                   (Xc - x0) * (Yc - y0) * inv};
    int id[4] = {M.ig[(iy - 1) + (ix - 1) * ny], M.ig[(iy - 1) + ix * ny], M.ig[iy + (ix - 1) * ny], M.ig[iy + ix * ny]};
    double s = 0;
    for (int c = 0; c < 4; ++c)
        if (w[c] != 0) s += w[c] * bicubic(M.grids + (size_t)id[c] * nn, M.gny, M.gnx, v, u, M.fill);
    return F * s;
}

bool make_model(const cp_psf_desc* d, HostModel& M) {
    if (!d) return false;
    M.kind = d->kind;
    M.fwhm = d->fwhm;
    M.alpha = d->alpha;
    M.beta = d->beta;
    if (d->kind == CP_PSF_CIRCULAR_GAUSSIAN) return d->fwhm > 0;
    if (d->kind == CP_PSF_CIRCULAR_MOFFAT) return d->alpha > 0 && d->beta > 1;
    if (d->kind != CP_PSF_GRIDDED_IMAGE || d->n_nodes < 1 || d->gny < 4 || d->gnx < 4 || !d->grids) return false;
    M.n_nodes = d->n_nodes;
    M.gny = d->gny;
    M.gnx = d->gnx;
    M.osx = d->os_x;
    M.osy = d->os_y;
    M.ox = d->origin_x;
    M.oy = d->origin_y;
    M.fill = d->fill_value;
    M.grids = d->grids;
    if (d->n_nodes > 1) {
        if (!d->node_y || !d->node_x) return false;
        M.yg.assign(d->node_y, d->node_y + d->n_nodes);
        M.xg.assign(d->node_x, d->node_x + d->n_nodes);
        std::sort(M.yg.begin(), M.yg.end());
        M.yg.erase(std::unique(M.yg.begin(), M.yg.end()), M.yg.end());
        std::sort(M.xg.begin(), M.xg.end());
        M.xg.erase(std::unique(M.xg.begin(), M.xg.end()), M.xg.end());
        int ny = (int)M.yg.size(), nx = (int)M.xg.size();
        if (ny < 2 || nx < 2 || ny * nx != d->n_nodes) return false;
        M.ig.assign((size_t)ny * nx, 0);
        for (int i = 0; i < d->n_nodes; ++i) {
            int a = (int)(std::lower_bound(M.yg.begin(), M.yg.end(), d->node_y[i]) - M.yg.begin());
            int b = (int)(std::lower_bound(M.xg.begin(), M.xg.end(), d->node_x[i]) - M.xg.begin());
            M.ig[a + b * ny] = i;
        }
    }
    return true;
}

}  // namespace

extern "C" {

int cp_sim_sources(uint64_t seed, int32_t n, double ylo, double yhi, double xlo, double xhi, int32_t flux_dist,
                   double flo, double fhi, double flux_power, double* y, double* x, double* flux) {
    if (n < 0 || !(flo > 0) || !(fhi >= flo) || flux_dist < 0 || flux_dist > 2) return CP_EARG;
    Rng r(seed);
    for (int i = 0; i < n; ++i) {
        y[i] = ylo + r.uniform() * (yhi - ylo);
        x[i] = xlo + r.uniform() * (xhi - xlo);
        flux[i] = sample_flux(r, flux_dist, flo, fhi, flux_power);
    }
    return CP_OK;
}

int cp_sim_sources_plummer(uint64_t seed, int32_t n, double a, double cy, double cx, double lo, double hi,
                           int32_t flux_dist, double flo, double fhi, double flux_power, double* y, double* x,
                           double* flux) {
    if (n < 0 || !(a > 0) || !(hi > lo) || !(flo > 0) || !(fhi >= flo)) return CP_EARG;
    Rng r(seed);
    int i = 0;
    while (i < n) {
        // projected Plummer: surface-density radius R = a sqrt(u / (1 - u))
        double u = r.uniform();
        double R = a * std::sqrt(u / (1 - u));
        double th = 2 * M_PI * r.uniform();
        double yy = cy + R * std::sin(th), xx = cx + R * std::cos(th);
        if (yy < lo || yy > hi || xx < lo || xx > hi) continue;
        y[i] = yy;
        x[i] = xx;
        flux[i] = sample_flux(r, flux_dist, flo, fhi, flux_power);
        ++i;
    }
    return CP_OK;
}

int cp_render_sources(const cp_psf_desc* psf, int32_t ny, int32_t nx, const double* y, const double* x,
                      const double* flux, int32_t n, double model_radius, double* image) {
    HostModel M;
    if (!make_model(psf, M) || ny <= 0 || nx <= 0 || n < 0 || !(model_radius > 0)) return CP_EARG;
    // Column strips, each rendered by one thread in source order (deterministic).
    const int strip = 64;
    const int nstrip = (nx + strip - 1) / strip;
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
#pragma omp parallel for schedule(dynamic, 1)
    for (int sidx = 0; sidx < nstrip; ++sidx) {
        const int c0 = sidx * strip + 1, c1 = std::min(nx, c0 + strip - 1);
        for (int i = 0; i < n; ++i) {
            if (!(std::isfinite(y[i]) && std::isfinite(x[i]) && std::isfinite(flux[i]))) continue;
            int ylo = std::max(1, (int)std::floor(y[i] - model_radius));
            int yhi = std::min(ny, (int)std::ceil(y[i] + model_radius));
            int xlo = std::max(c0, (int)std::floor(x[i] - model_radius));
            int xhi = std::min(c1, (int)std::ceil(x[i] + model_radius));
            for (int gx = xlo; gx <= xhi; ++gx)
                for (int gy = ylo; gy <= yhi; ++gy)
                    image[(size_t)(gy - 1) + (size_t)(gx - 1) * ny] += eval_model(M, y[i], x[i], flux[i], gy, gx);
        }
    }
    return CP_OK;
}

int cp_add_noise(uint64_t seed, double* image, int64_t npix, double gain, double read_noise) {
    if (!(gain > 0) || read_noise < 0 || npix < 0) return CP_EARG;
    const int64_t blk = 65536;
    const int64_t nb = (npix + blk - 1) / blk;
#pragma omp parallel for schedule(static)
    for (int64_t b = 0; b < nb; ++b) {
        Rng r(seed * 0x100000001b3ULL + (uint64_t)b * 0x9e3779b97f4a7c15ULL + 1);
        int64_t q0 = b * blk, q1 = std::min(npix, q0 + blk);
        for (int64_t q = q0; q < q1; ++q) image[q] = (double)rand_poisson(r, std::max(0.0, image[q] * gain)) / gain;
        if (read_noise > 0)
            for (int64_t q = q0; q < q1; ++q) image[q] += read_noise * r.normal();
    }
    return CP_OK;
}

}  // extern "C"
