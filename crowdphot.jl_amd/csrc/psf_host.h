// psf_host.h -- host-side PSF descriptor handling shared by the HIP
// translation units: validation (the ArgumentError contract of the model
// constructors, src/psf/gridded_psf.jl:68-144, src/psf/empirical_models.jl:91-140),
// upload of ePSF node grids, and the per-model device constants.
#pragma once

#include <algorithm>
#include <cmath>
#include <vector>

#include "sim_kernels.h"

namespace cp {

// Kernel launch; with profiling on, bracket it with HIP events on its stream.
template <typename K, typename... Args>
static inline void launch_n(const char* name, K kernel, dim3 grid, dim3 block, size_t shmem, hipStream_t s,
                            Args... args) {
    cp_ctx* c = tl_ctx;
    hipEvent_t a = nullptr, b = nullptr;
    if (c && c->prof) {
        for (hipEvent_t* e : {&a, &b}) {
            if (!c->ev_pool.empty()) {
                *e = c->ev_pool.back();
                c->ev_pool.pop_back();
            } else {
                CP_HIP(hipEventCreate(e));
            }
        }
        CP_HIP(hipEventRecord(a, s));
    }
    hipLaunchKernelGGL(kernel, grid, block, shmem, s, args...);
    CP_HIP(hipGetLastError());
    if (a) {
        CP_HIP(hipEventRecord(b, s));
        c->prof_pending.push_back({name, a, b});
        if (c->prof_pending.size() > 4096) prof_flush(c);
    }
}

template <typename K, typename... Args>
static inline void launch(K kernel, dim3 grid, dim3 block, size_t shmem, hipStream_t s, Args... args) {
    launch_n("other", kernel, grid, block, shmem, s, args...);
}

static inline unsigned int cdiv(int64_t a, int64_t b) { return (unsigned int)((a + b - 1) / b); }

// ---------------------------------------------------------------------------
// PSF descriptor -> host + device representation
// ---------------------------------------------------------------------------
struct HostPsf {
    int kind = 0;
    double fwhm = 0, alpha = 0, beta = 0;
    int n_nodes = 0, gny = 0, gnx = 0;
    double osx = 1, osy = 1, ox = 0, oy = 0, fill = 0;
    std::vector<double> ygrid, xgrid;
    std::vector<int> index_grid;
};

inline HostPsf make_host_psf(const cp_psf_desc* d) {
    CP_ARG(d != nullptr, "psf descriptor is NULL");
    HostPsf P;
    P.kind = d->kind;
    if (d->kind == CP_PSF_CIRCULAR_GAUSSIAN) {
        CP_ARG(d->fwhm > 0 && std::isfinite(d->fwhm), "fwhm must be positive and finite");
        P.fwhm = d->fwhm;
        return P;
    }
    if (d->kind == CP_PSF_CIRCULAR_MOFFAT) {
        CP_ARG(d->alpha > 0 && d->beta > 1, "Moffat needs alpha > 0 and beta > 1");
        P.alpha = d->alpha;
        P.beta = d->beta;
        return P;
    }
    CP_ARG(d->kind == CP_PSF_GRIDDED_IMAGE, "unknown psf kind");
    CP_ARG(d->n_nodes >= 1, "`psfs` must be non-empty");
    CP_ARG(d->gny >= 4 && d->gnx >= 4, "each axis of `data` must have at least 4 samples");
    CP_ARG(d->grids != nullptr, "ePSF grids pointer is NULL");
    CP_ARG(d->os_x >= 1 && d->os_y >= 1, "oversampling must be a positive integer");
    P.n_nodes = d->n_nodes;
    P.gny = d->gny;
    P.gnx = d->gnx;
    P.osx = d->os_x;
    P.osy = d->os_y;
    P.ox = d->origin_x;
    P.oy = d->origin_y;
    P.fill = d->fill_value;
    if (d->n_nodes == 1) {
        P.index_grid = {0};
        return P;
    }
    CP_ARG(d->node_y && d->node_x, "node positions are required for more than one node");
    std::vector<double> gy(d->node_y, d->node_y + d->n_nodes), gx(d->node_x, d->node_x + d->n_nodes);
    P.ygrid = gy;
    P.xgrid = gx;
    std::sort(P.ygrid.begin(), P.ygrid.end());
    P.ygrid.erase(std::unique(P.ygrid.begin(), P.ygrid.end()), P.ygrid.end());
    std::sort(P.xgrid.begin(), P.xgrid.end());
    P.xgrid.erase(std::unique(P.xgrid.begin(), P.xgrid.end()), P.xgrid.end());
    int ngy = (int)P.ygrid.size(), ngx = (int)P.xgrid.size();
    CP_ARG(ngy >= 2 && ngx >= 2, "`grid_y` and `grid_x` must each contain at least 2 unique values");
    CP_ARG(ngy * ngx == d->n_nodes, "node positions do not form a complete rectangular grid");
    P.index_grid.assign((size_t)ngy * ngx, -1);
    for (int i = 0; i < d->n_nodes; ++i) {
        int iy = (int)(std::lower_bound(P.ygrid.begin(), P.ygrid.end(), gy[i]) - P.ygrid.begin());
        int ix = (int)(std::lower_bound(P.xgrid.begin(), P.xgrid.end(), gx[i]) - P.xgrid.begin());
        int& slot = P.index_grid[(size_t)iy + (size_t)ix * ngy];
        CP_ARG(slot < 0, "duplicate node position");
        slot = i;
    }
    return P;
}

struct DevPsf {
    DevBuf grids, ygrid, xgrid, index_grid;
    NodeGrid G{};
};

inline void upload_psf(cp_ctx* ctx, const cp_psf_desc* d, const HostPsf& P, DevPsf& D) {
    if (P.kind != CP_PSF_GRIDDED_IMAGE) return;
    size_t ng = (size_t)P.n_nodes * P.gny * P.gnx;
    D.grids.alloc(ctx, ng * sizeof(float));
    CP_HIP(hipMemcpyAsync(D.grids.get(), d->grids, ng * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    D.G.ngy = (int)P.ygrid.size();
    D.G.ngx = (int)P.xgrid.size();
    if (P.n_nodes > 1) {
        D.ygrid.alloc(ctx, P.ygrid.size() * 8);
        D.xgrid.alloc(ctx, P.xgrid.size() * 8);
        D.index_grid.alloc(ctx, P.index_grid.size() * 4);
        CP_HIP(hipMemcpyAsync(D.ygrid.get(), P.ygrid.data(), P.ygrid.size() * 8, hipMemcpyHostToDevice, ctx->stream));
        CP_HIP(hipMemcpyAsync(D.xgrid.get(), P.xgrid.data(), P.xgrid.size() * 8, hipMemcpyHostToDevice, ctx->stream));
        CP_HIP(hipMemcpyAsync(D.index_grid.get(), P.index_grid.data(), P.index_grid.size() * 4,
                              hipMemcpyHostToDevice, ctx->stream));
        D.G.ygrid = D.ygrid.as<double>();
        D.G.xgrid = D.xgrid.as<double>();
        D.G.index_grid = D.index_grid.as<int>();
    }
}

// Model constants over node grids already in device memory (grids may be
// null for the analytic models).
template <typename T>
ModelConst<T> model_const(const HostPsf& P, const float* grids, double bkg) {
    ModelConst<T> M{};
    const double GP = CP_GAUSS_PRE;
    if (P.kind == CP_PSF_CIRCULAR_GAUSSIAN) {
        double f2 = P.fwhm * P.fwhm;
        double norm = -(M_PI * f2 / GP);
        M.fwhm2 = (T)f2;
        M.inv_fwhm2 = (T)(1.0 / f2);
        M.inv_norm_g = (T)(1.0 / norm);
    } else if (P.kind == CP_PSF_CIRCULAR_MOFFAT) {
        M.alpha2 = (T)(P.alpha * P.alpha);
        M.inv_alpha2 = (T)(1.0 / (P.alpha * P.alpha));
        M.beta = (T)P.beta;
        M.norm_m = (T)(M_PI * P.alpha * P.alpha / (P.beta - 1));
        M.inv_norm_m = (T)((P.beta - 1) / (M_PI * P.alpha * P.alpha));
    } else {
        M.gny = P.gny;
        M.gnx = P.gnx;
        M.n_nodes = P.n_nodes;
        M.osx = (T)P.osx;
        M.osy = (T)P.osy;
        M.ox = (T)P.ox;
        M.oy = (T)P.oy;
        M.fill = (T)P.fill;
        M.grids = grids;
        M.ios_y = (int)P.osy;
        M.ios_x = (int)P.osx;
        M.Qy = (P.gny + 2 * EPSF_PAD + M.ios_y - 1) / M.ios_y;
        M.Qx = (P.gnx + 2 * EPSF_PAD + M.ios_x - 1) / M.ios_x;
    }
    M.bkg = (T)bkg;
    return M;
}

template <typename T>
ModelConst<T> model_const(const HostPsf& P, const DevPsf& D, double bkg) {
    return model_const<T>(P, D.grids.as<float>(), bkg);
}

}  // namespace cp
