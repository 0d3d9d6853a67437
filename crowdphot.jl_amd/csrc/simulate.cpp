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
#include "host_model.h"

namespace {
using namespace cp_host;

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
