// context.hip -- context lifetime, defaults and the model-evaluation entry point.
#include <algorithm>
#include <cstring>

#include "sim_kernels.h"
#include "psf_host.h"

namespace cp {

// f and the full getproperties-ordered gradient of one model at n points
// (evaluate_fg, src/psf/PSF.jl:207-226), float64, through the same device
// formulas the fit kernels use.
template <int KIND>
__global__ void k_evaluate(ModelConst<double> M, NodeGrid G, double Y, double X, double F, double fwhm,
                           double alpha, double beta, const double* __restrict__ py,
                           const double* __restrict__ px, int64_t n, double* __restrict__ f,
                           double* __restrict__ Gout) {
    int64_t i = blockIdx.x * (int64_t)BLOCK + threadIdx.x;
    if (i >= n) return;
    StarRec<double> r{};
    r.flux = F;
    if (KIND == KIND_GAUSS) r.amp = F * M.inv_norm_g;
    else if (KIND == KIND_MOFFAT) r.amp = F * (beta - 1) / (M_PI * alpha * alpha);
    else r.amp = F;
    CornerRec<double> cr;
    if (KIND == KIND_EPSF) grid_corners<double>(G, M.n_nodes, Y, X, cr);
    const double ddy = py[i] - Y, ddx = px[i] - X;
    double gy, gx, gf;
    f[i] = star_value<double, KIND>(M, r, ddy, ddx, M.grids, &cr);
    star_grad<double, KIND>(M, r, ddy, ddx, M.grids, &cr, gy, gx, gf);
    if (KIND == KIND_GAUSS) {
        double sq = (ddx * ddx + ddy * ddy) / M.fwhm2;
        double Ag = r.amp * exp(CP_GAUSS_PRE * sq);
        double* g = Gout + i * 5;
        g[0] = gy; g[1] = gx; g[2] = -2 * Ag * (1 + CP_GAUSS_PRE * sq) / fwhm; g[3] = gf; g[4] = 1.0;
    } else if (KIND == KIND_MOFFAT) {
        double r2 = ddx * ddx + ddy * ddy, u = 1 + r2 / M.alpha2;
        double Ag = (F / M.norm_m) * pow(u, -beta);
        double* g = Gout + i * 6;
        g[0] = gy; g[1] = gx;
        g[2] = Ag * (-2 / alpha + 2 * beta * r2 / (alpha * M.alpha2 * u));
        g[3] = Ag * (1 / (beta - 1) - log(u));
        g[4] = gf; g[5] = 1.0;
    } else {
        double* g = Gout + i * 4;
        g[0] = gy; g[1] = gx; g[2] = gf; g[3] = 1.0;
    }
}

thread_local cp_ctx* tl_ctx = nullptr;

void staged_d2h(cp_ctx* c, hipStream_t st, const std::vector<D2HSeg>& segs) {
    // chunks of at most STAGE_CHUNK bytes, each one contiguous run of device
    // memory feeding one or more host destinations
    struct Piece {
        char* dst;
        size_t off, bytes;   // offset inside the chunk
    };
    struct Chunk {
        const char* src;
        size_t bytes;
        std::vector<Piece> pieces;
    };
    std::vector<Chunk> ch;
    for (const D2HSeg& sg : segs) {
        const char* src = static_cast<const char*>(sg.src);
        char* dst = static_cast<char*>(sg.dst);
        size_t left = sg.bytes;
        while (left > 0) {
            if (ch.empty() || ch.back().src + ch.back().bytes != src || ch.back().bytes == STAGE_CHUNK)
                ch.push_back(Chunk{src, 0, {}});
            Chunk& k = ch.back();
            const size_t take = std::min(left, STAGE_CHUNK - k.bytes);
            k.pieces.push_back(Piece{dst, k.bytes, take});
            k.bytes += take;
            src += take;
            dst += take;
            left -= take;
        }
    }
    if (ch.empty()) return;
    if (!c->stage) {
        CP_HIP(hipHostMalloc(&c->stage, 2 * STAGE_CHUNK, hipHostMallocDefault));
        for (auto& e : c->stage_ev) CP_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    char* buf[2] = {static_cast<char*>(c->stage), static_cast<char*>(c->stage) + STAGE_CHUNK};
    auto issue = [&](size_t k) {
        CP_HIP(hipMemcpyAsync(buf[k & 1], ch[k].src, ch[k].bytes, hipMemcpyDeviceToHost, st));
        CP_HIP(hipEventRecord(c->stage_ev[k & 1], st));
    };
    issue(0);
    for (size_t k = 0; k < ch.size(); ++k) {
        // buffer (k + 1) & 1 last held chunk k - 1, already copied out
        if (k + 1 < ch.size()) issue(k + 1);
        CP_HIP(hipEventSynchronize(c->stage_ev[k & 1]));
        for (const Piece& pc : ch[k].pieces) par_memcpy(pc.dst, buf[k & 1] + pc.off, pc.bytes);
    }
}

}  // namespace cp

using namespace cp;

extern "C" {

int cp_create(cp_ctx** out, int device, uint64_t hbm_budget_bytes) {
    if (!out) return CP_EARG;
    *out = nullptr;
    cp_ctx* c = new (std::nothrow) cp_ctx();
    if (!c) return CP_ENOMEM;
    int rc = guarded(c, [&] {
        int ndev = 0;
        CP_HIP(hipGetDeviceCount(&ndev));
        CP_ARG(device >= 0 && device < ndev, "no such HIP device");
        CP_HIP(hipSetDevice(device));
        c->device = device;
        c->budget = hbm_budget_bytes;
        CP_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        CP_HIP(hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, device));
    });
    if (rc != CP_OK) {
        delete c;
        return rc;
    }
    *out = c;
    return CP_OK;
}

void cp_destroy(cp_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) {
        (void)hipStreamSynchronize(c->stream);
        (void)hipStreamDestroy(c->stream);
    }
    if (c->pinned) (void)hipHostFree(c->pinned);
    if (c->stage) (void)hipHostFree(c->stage);
    for (auto e : c->stage_ev)
        if (e) (void)hipEventDestroy(e);
    cp::cache_trim(c, UINT64_MAX);
    for (auto& p : c->prof_pending) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    for (auto e : c->ev_pool) (void)hipEventDestroy(e);
    delete c;
}

const char* cp_last_error(const cp_ctx* c) { return c ? c->err.c_str() : "NULL context"; }

void* cp_stream(cp_ctx* c) { return c ? (void*)c->stream : nullptr; }

int cp_synchronize(cp_ctx* c) {
    if (!c) return CP_ESTATE;
    return guarded(c, [&] { CP_HIP(hipStreamSynchronize(c->stream)); });
}

int cp_get_tuning(cp_ctx* c, cp_tuning* out) {
    if (!c) return CP_ESTATE;
    return guarded(c, [&] {
        CP_ARG(out != nullptr, "NULL argument");
        *out = c->tuning;
    });
}

int cp_set_tuning(cp_ctx* c, const cp_tuning* t) {
    if (!c) return CP_ESTATE;
    return guarded(c, [&] {
        CP_ARG(t != nullptr, "NULL argument");
        for (int32_t v : {t->hbuild_blocks_per_cu, t->cg_blocks_per_cu, t->render_blocks_per_cu, t->lm_blocks_per_cu})
            CP_ARG(v >= 0 && v <= 64, "blocks per CU must be in [0, 64] (0 = default), got " + std::to_string(v));
        CP_ARG(t->render_tile == 0 || t->render_tile == 32 || t->render_tile == 64,
               "render_tile must be 0 (default), 32 or 64, got " + std::to_string(t->render_tile));
        CP_ARG(t->hbuild_lanes == 0 || t->hbuild_lanes == 4 || t->hbuild_lanes == 16 || t->hbuild_lanes == 64,
               "hbuild_lanes must be 0 (default), 4, 16 or 64, got " + std::to_string(t->hbuild_lanes));
        CP_ARG(t->fill_lanes == 0 || t->fill_lanes == 16 || t->fill_lanes == 64,
               "fill_lanes must be 0 (default), 16 or 64, got " + std::to_string(t->fill_lanes));
        CP_ARG(t->lm_work_tile >= 0 && t->lm_work_tile <= 64,
               "lm_work_tile must be in [0, 64] (0 = default), got " + std::to_string(t->lm_work_tile));
        CP_ARG(t->gk_fold_blocks >= 0 && t->gk_fold_blocks <= (1 << 20),
               "gk_fold_blocks must be in [0, 1048576] (0 = default), got " + std::to_string(t->gk_fold_blocks));
        c->tuning = *t;
    });
}

int cp_mem_stats_get(cp_ctx* c, cp_mem_stats* out) {
    if (!c) return CP_ESTATE;
    return guarded(c, [&] {
        CP_ARG(out != nullptr, "NULL argument");
        out->budget = c->budget;
        out->in_use = c->in_use;
        out->cached = c->cached;
        out->peak = c->peak;
        out->peak_resident = c->peak_resident;
    });
}

int cp_set_budget(cp_ctx* c, uint64_t budget) {
    if (!c) return CP_ESTATE;
    return guarded(c, [&] {
        CP_HIP(hipSetDevice(c->device));
        c->budget = budget;
        // cached blocks beyond the new cap go back to the device; the peak
        // restarts from what is in use now
        if (budget && c->in_use + c->cached > budget) {
            CP_HIP(hipStreamSynchronize(c->stream));
            cache_trim(c, 0);
        }
        c->peak = c->in_use;
        c->peak_resident = c->in_use + c->cached;
    });
}

int cp_release_cached(cp_ctx* c) {
    if (!c) return CP_ESTATE;
    return guarded(c, [&] {
        CP_HIP(hipSetDevice(c->device));
        CP_HIP(hipStreamSynchronize(c->stream));
        cache_trim(c, UINT64_MAX);
    });
}

int cp_profile(cp_ctx* c, int enable) {
    if (!c) return CP_ESTATE;
    return guarded(c, [&] {
        CP_HIP(hipStreamSynchronize(c->stream));
        prof_flush(c);
        c->prof = enable != 0;
        if (enable) c->prof_acc.clear();
    });
}

int cp_profile_report(cp_ctx* c, char* buf, int32_t cap) {
    if (!c || !buf || cap <= 0) return CP_ESTATE;
    return guarded(c, [&] {
        prof_flush(c);
        std::string s;
        char line[256];
        for (auto& kv : c->prof_acc) {
            snprintf(line, sizeof line, "%s %ld %.6f\n", kv.first.c_str(), kv.second.second, kv.second.first);
            s += line;
        }
        snprintf(buf, cap, "%s", s.c_str());
    });
}

int cp_psf_nparams(int32_t kind) {
    switch (kind) {
        case CP_PSF_CIRCULAR_GAUSSIAN: return 5;
        case CP_PSF_CIRCULAR_MOFFAT: return 6;
        case CP_PSF_GRIDDED_IMAGE: return 4;
    }
    return CP_EARG;
}

int cp_evaluate_fg(cp_ctx* ctx, const cp_psf_desc* psf, const double* prm, const double* py, const double* px,
                   int64_t n, double* f, double* G) {
    if (!ctx) return CP_ESTATE;
    return guarded(ctx, [&] {
        CP_ARG(prm && py && px && f && G && n >= 0, "NULL argument");
        HostPsf P = make_host_psf(psf);
        CP_HIP(hipSetDevice(ctx->device));
        hipStream_t st = ctx->stream;
        DevPsf DP;
        upload_psf(ctx, psf, P, DP);
        const int np = cp_psf_nparams(P.kind);
        if (n == 0) return;
        DevBuf dy(ctx, n * 8), dx(ctx, n * 8), df(ctx, n * 8), dG(ctx, n * np * 8);
        CP_HIP(hipMemcpyAsync(dy.get(), py, n * 8, hipMemcpyHostToDevice, st));
        CP_HIP(hipMemcpyAsync(dx.get(), px, n * 8, hipMemcpyHostToDevice, st));
        double bkg = prm[np - 1], F = prm[np - 2];
        ModelConst<double> M = model_const<double>(P, DP, bkg);
        dim3 g((unsigned)((n + BLOCK - 1) / BLOCK));
        if (P.kind == CP_PSF_CIRCULAR_GAUSSIAN) {
            CP_ARG(prm[2] > 0, "fwhm must be positive");
            double f2 = prm[2] * prm[2];
            M.fwhm2 = f2;
            M.inv_fwhm2 = 1.0 / f2;
            M.inv_norm_g = 1.0 / (-(M_PI * f2 / CP_GAUSS_PRE));
            hipLaunchKernelGGL(k_evaluate<KIND_GAUSS>, g, dim3(BLOCK), 0, st, M, DP.G, prm[0], prm[1], F, prm[2],
                               0.0, 0.0, dy.as<double>(), dx.as<double>(), n, df.as<double>(), dG.as<double>());
        } else if (P.kind == CP_PSF_CIRCULAR_MOFFAT) {
            double a = prm[2], b = prm[3];
            M.alpha2 = a * a;
            M.inv_alpha2 = 1.0 / (a * a);
            M.beta = b;
            M.norm_m = M_PI * a * a / (b - 1);
            M.inv_norm_m = (b - 1) / (M_PI * a * a);
            hipLaunchKernelGGL(k_evaluate<KIND_MOFFAT>, g, dim3(BLOCK), 0, st, M, DP.G, prm[0], prm[1], F, 0.0, a,
                               b, dy.as<double>(), dx.as<double>(), n, df.as<double>(), dG.as<double>());
        } else {
            hipLaunchKernelGGL(k_evaluate<KIND_EPSF>, g, dim3(BLOCK), 0, st, M, DP.G, prm[0], prm[1], F, 0.0, 0.0,
                               0.0, dy.as<double>(), dx.as<double>(), n, df.as<double>(), dG.as<double>());
        }
        CP_HIP(hipGetLastError());
        CP_HIP(hipMemcpyAsync(f, df.get(), n * 8, hipMemcpyDeviceToHost, st));
        CP_HIP(hipMemcpyAsync(G, dG.get(), n * np * 8, hipMemcpyDeviceToHost, st));
        CP_HIP(hipStreamSynchronize(st));
    });
}

}  // extern "C"
