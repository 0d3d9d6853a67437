// host_model.h -- float64 host evaluation of the PSF models (the same formulas
// as the device models): CircularGaussianPSF (src/psf/parametric_models.jl:62-113),
// CircularMoffatPSF (:672-724), ImagePSF bicubic (src/psf/empirical_models.jl:185-318)
// and the GriddedPSFModel bilinear node blend (src/psf/gridded_psf.jl:176-271).
// Used by the input generator and the one-shot diagnostics set-up.
#pragma once

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/crowdphot_hip.h"

namespace cp_host {

inline const double GAUSS_PRE = -4.0 * std::log(2.0);

struct HostModel {
    int kind;
    double fwhm, alpha, beta;
    int n_nodes, gny, gnx;
    double osx, osy, ox, oy, fill;
    const float* grids;
    std::vector<double> yg, xg;
    std::vector<int> ig;
};

inline void cubic4v(double f1, double f2, double f3, double f4, double t, double& v) {
    double c1 = (f3 - f1) / 2, c4 = f3 - f2 - c1, c2 = 3 * c4 - (f4 - f2) / 2 + c1, c3 = c4 - c2;
    v = t * (t * (t * c3 + c2) + c1) + f2;
}

inline double bicubic(const float* g, int ny, int nx, double y, double x, double fill) {
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

inline double eval_model(const HostModel& M, double Y, double X, double F, double py, double px) {
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

inline bool make_model(const cp_psf_desc* d, HostModel& M) {
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

}  // namespace cp_host
