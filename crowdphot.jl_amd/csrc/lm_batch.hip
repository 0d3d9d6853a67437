// lm_batch.hip -- batched per-star Levenberg-Marquardt (fit_star over many
// uneven cutouts) on gfx950.
//
// One wavefront owns one cutout for its whole fit (lm_fit); the cutout is read
// once from HBM and stays L1/L2-resident across the accum! passes.  Waves pull
// cutouts from an integer work counter until it passes the end.
//
// References: fit_star src/psf/psf_fitting.jl:229-308; lm_irls
// src/levenberg_marquardt.jl:431-680.
#include <cstring>
#include <mutex>
#include <type_traits>

#include "lm_core.h"
#include "psf_host.h"

namespace cp {

// Waves per SIMD of a batched-LM instance: three for the float32 fits of an
// analytic model with up to four free parameters (168 registers; the spills
// this costs sit in the per-fit setup, the damped step and the result write,
// not in the per-pixel passes: C3 18.2 -> 15.4 ms), the compiler's choice
// otherwise (the ePSF fits would spill inside their bicubic passes).
template <typename T, int KIND, int NP>
constexpr int lm_waves() {
    return sizeof(T) == 4 && KIND != KIND_EPSF && NP <= 4 ? 3 : 1;
}

template <typename T, int KIND, int NP, bool UNIT_W>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(lm_waves<T, KIND, NP>()))) void k_lm_batch(LMConst C, LMModel<T> Md, const float* __restrict__ pix,
                                                    const int64_t* __restrict__ offsets,
                                                    const int32_t* __restrict__ cny, const int32_t* __restrict__ cnx,
                                                    const float* __restrict__ iv, const double* __restrict__ theta0,
                                                    int n, double* __restrict__ theta_out,
                                                    double* __restrict__ cov_out, cp_lm_stat* __restrict__ stats,
                                                    int* __restrict__ work, int wtile) {
    const int lane = threadIdx.x & 63;
    for (;;) {
        // a work tile of wtile consecutive cutouts per counter increment
        int c0 = 0;
        if (lane == 0) c0 = atomicAdd(work, wtile);
        c0 = __shfl(c0, 0, 64);
        if (c0 >= n) return;
        const int c1 = min(c0 + wtile, n);
        for (int c = c0; c < c1; ++c) {
            const int ny = cny[c], nx = cnx[c];
            const int64_t off = offsets[c];
            double full0[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
            for (int k = 0; k < 6; ++k)
                if (k < C.npf) full0[k] = theta0[(size_t)c * C.npf + k];
            LMOut<NP> R;
            if constexpr (UNIT_W) {
                const CutoutSrcUnit S{pix + off, ny, ny * nx};
                lm_fit<T, KIND, NP>(C, Md, full0, S, R);
            } else {
                const CutoutSrc S{pix + off, C.have_iv ? iv + off : nullptr, ny, ny * nx};
                lm_fit<T, KIND, NP>(C, Md, full0, S, R);
            }
            // Lane 0 writes inside an if that falls through to the loop end (no
            // divergent `continue`): the loops then have one back edge each and
            // the whole wave is converged again when the next __shfl reads lane 0.
            if (lane == 0) {
                // (a constant-bound loop: an index bounded by C.npf would put R in scratch)
#pragma unroll
                for (int k = 0; k < 6; ++k)
                    if (k < C.npf) theta_out[(size_t)c * C.npf + k] = R.full[k];
                if (R.st.status == 0 && cov_out && C.cov_mode != CP_COV_NONE) {
                    double cv[NP * NP];
                    lm_cov<NP>(C, R, ny * nx, cv);
                    for (int i = 0; i < NP * NP; ++i) cov_out[(size_t)c * NP * NP + i] = cv[i];
                }
                stats[c] = R.st;
            }
        }
    }
}

template <typename T, int KIND>
void launch_lm(int np, dim3 g, hipStream_t s, const LMConst& C, const LMModel<T>& Md, const float* pix,
               const int64_t* off, const int32_t* ny, const int32_t* nx, const float* iv, const double* th0, int n,
               double* th, double* cov, cp_lm_stat* st, int* work, int wtile) {
#define CP_LM(NPV)                                                                                          \
    (C.have_iv ? launch_n("k_lm_batch", k_lm_batch<T, KIND, NPV, false>, g, dim3(BLOCK), 0, s, C, Md, pix, off, ny, \
                          nx, iv, th0, n, th, cov, st, work, wtile)                                            \
               : launch_n("k_lm_batch", k_lm_batch<T, KIND, NPV, true>, g, dim3(BLOCK), 0, s, C, Md, pix, off, ny,  \
                          nx, iv, th0, n, th, cov, st, work, wtile))
    switch (np) {
        case 1: CP_LM(1); break;
        case 2: CP_LM(2); break;
        case 3: CP_LM(3); break;
        case 4: CP_LM(4); break;
        case 5: CP_LM(5); break;
        default: CP_LM(6); break;
    }
#undef CP_LM
}

// Launches of one batched call; the work counter and the ePSF node grids are
// already in device memory (G and grids unused for the analytic models).
void lm_batch_launch(cp_ctx* ctx, const cp_psf_desc* psf, const HostPsf& P, const LMConst& C, int np,
                     const float* grids, const NodeGrid& G, int* work, const float* d_pix, const int64_t* d_off,
                     const int32_t* d_ny, const int32_t* d_nx, const float* d_iv, const double* d_th0, int n,
                     double* d_th, double* d_cov, cp_lm_stat* d_st, int precision) {
    hipStream_t st = ctx->stream;
    dim3 g((unsigned)std::min<int64_t>((n + NWAVE - 1) / NWAVE,
                                       (int64_t)std::max(ctx->n_cu, 1) * tuned_bpc(ctx->tuning.lm_blocks_per_cu, 8)));
    auto go = [&](auto tag, auto kind) {
        using T = decltype(tag);
        constexpr int K = decltype(kind)::value;
        const LMModel<T> Md{model_const<T>(P, grids, 0.0), G};
        launch_lm<T, K>(np, g, st, C, Md, d_pix, d_off, d_ny, d_nx, d_iv, d_th0, n, d_th, d_cov, d_st, work,
                        ctx->tuning.lm_work_tile > 0 ? ctx->tuning.lm_work_tile : 1);
    };
    using KG = std::integral_constant<int, KIND_GAUSS>;
    using KM = std::integral_constant<int, KIND_MOFFAT>;
    using KE = std::integral_constant<int, KIND_EPSF>;
    const bool f64 = precision == CP_PREC_FP64;
    if (psf->kind == CP_PSF_CIRCULAR_GAUSSIAN) f64 ? go(0.0, KG{}) : go(0.0f, KG{});
    else if (psf->kind == CP_PSF_CIRCULAR_MOFFAT) f64 ? go(0.0, KM{}) : go(0.0f, KM{});
    else f64 ? go(0.0, KE{}) : go(0.0f, KE{});
    CP_HIP(hipStreamSynchronize(st));
}

void lm_batch_device(cp_ctx* ctx, const cp_psf_desc* psf, const float* d_pix, const int64_t* d_off,
                     const int32_t* d_ny, const int32_t* d_nx, const float* d_iv, const double* d_th0,
                     uint32_t free_mask, int n, const cp_lm_opts* o, double* d_th, double* d_cov,
                     cp_lm_stat* d_st, int precision) {
    CP_ARG(psf && o, "NULL argument");
    LMConst C;
    int np = 0;
    const char* why = "";
    CP_ARG(lm_const(psf, free_mask, o, d_iv != nullptr, C, np, &why), why);
    if (n == 0) return;
    const HostPsf P = make_host_psf(psf);
    DevPsf DP;
    upload_psf(ctx, psf, P, DP);
    DevBuf work(ctx, 4);
    CP_HIP(hipMemsetAsync(work.get(), 0, 4, ctx->stream));
    lm_batch_launch(ctx, psf, P, C, np, DP.grids.as<float>(), DP.G, work.as<int>(), d_pix, d_off, d_ny, d_nx, d_iv,
                    d_th0, n, d_th, d_cov, d_st, precision);
}

// ---------------------------------------------------------------------------
// The allocation-free variant: the work counter and the ePSF node grids live
// in one file-scope device array per GPU instead of context blocks.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[CP_NOALLOC_SCRATCH_BYTES];

// scratch's device address on each GPU (a host-side use of the symbol itself
// is not a device address), looked up once per device with it current.
static std::mutex g_scratch_mu;
static std::vector<unsigned char*> g_scratch_ptr;

static unsigned char* scratch_on(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize(device + 1, nullptr);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        CP_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[device] = static_cast<unsigned char*>(p);
    }
    return g_scratch_ptr[device];
}

static __global__ void k_zero_int(int* p) {
    if (threadIdx.x == 0) *p = 0;
}

static size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }

void lm_batch_noalloc(cp_ctx* ctx, const cp_psf_desc* psf, const float* d_pix, const int64_t* d_off,
                      const int32_t* d_ny, const int32_t* d_nx, const float* d_iv, const double* d_th0,
                      uint32_t free_mask, int n, const cp_lm_opts* o, double* d_th, double* d_cov,
                      cp_lm_stat* d_st) {
    CP_ARG(psf && o, "NULL argument");
    LMConst C;
    int np = 0;
    const char* why = "";
    CP_ARG(lm_const(psf, free_mask, o, d_iv != nullptr, C, np, &why), why);
    const HostPsf P = make_host_psf(psf);
    // layout: [work counter | node grids | grid_y | grid_x | index_grid], 256-byte aligned pieces
    const bool epsf = P.kind == CP_PSF_GRIDDED_IMAGE;
    const bool multi = epsf && P.n_nodes > 1;
    const size_t ng = epsf ? (size_t)P.n_nodes * P.gny * P.gnx : 0;
    const size_t o_grids = 256, o_y = o_grids + round256(ng * 4), o_x = o_y + round256(multi ? P.ygrid.size() * 8 : 0),
                 o_idx = o_x + round256(multi ? P.xgrid.size() * 8 : 0),
                 end = o_idx + round256(multi ? P.index_grid.size() * 4 : 0);
    if (end > CP_NOALLOC_SCRATCH_BYTES)
        throw Error(CP_ENOMEM, "cp_fit_star_batch_device_noalloc: the PSF needs " + std::to_string(end) +
                                   " bytes of device scratch, more than its " +
                                   std::to_string((size_t)CP_NOALLOC_SCRATCH_BYTES) + "-byte maximum");
    if (n == 0) return;
    hipStream_t st = ctx->stream;
    unsigned char* base = scratch_on(ctx->device);
    int* work = reinterpret_cast<int*>(base);
    // a fill kernel, not a memset node, so the call can sit inside a captured graph
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, st, work);
    CP_HIP(hipGetLastError());
    NodeGrid G{};
    const float* grids = nullptr;
    if (epsf) {
        CP_HIP(hipMemcpyAsync(base + o_grids, psf->grids, ng * 4, hipMemcpyHostToDevice, st));
        grids = reinterpret_cast<const float*>(base + o_grids);
        G.ngy = (int)P.ygrid.size();
        G.ngx = (int)P.xgrid.size();
        if (multi) {
            CP_HIP(hipMemcpyAsync(base + o_y, P.ygrid.data(), P.ygrid.size() * 8, hipMemcpyHostToDevice, st));
            CP_HIP(hipMemcpyAsync(base + o_x, P.xgrid.data(), P.xgrid.size() * 8, hipMemcpyHostToDevice, st));
            CP_HIP(hipMemcpyAsync(base + o_idx, P.index_grid.data(), P.index_grid.size() * 4, hipMemcpyHostToDevice,
                                  st));
            G.ygrid = reinterpret_cast<const double*>(base + o_y);
            G.xgrid = reinterpret_cast<const double*>(base + o_x);
            G.index_grid = reinterpret_cast<const int*>(base + o_idx);
        }
    }
    lm_batch_launch(ctx, psf, P, C, np, grids, G, work, d_pix, d_off, d_ny, d_nx, d_iv, d_th0, n, d_th, d_cov, d_st,
                    CP_PREC_MIXED);
}

}  // namespace cp

using namespace cp;

extern "C" {

void cp_lm_opts_default(cp_lm_opts* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->lambda_init = 1e-4;
    o->lambda_up = 10.0;
    o->lambda_down = 10.0;
    o->lambda_min = 1e-12;
    o->lambda_max = 1e12;
    o->max_iter = 200;
    o->x_tol = 1e-8;
    o->f_tol = 1e-4;
    o->g_tol = 1e-4;
    o->damping = CP_DAMP_MARQUARDT;
    o->min_diagonal = 1e-6;
    o->covariance = CP_COV_DEFAULT;
}

int cp_fit_star_batch_device(cp_ctx* ctx, const cp_psf_desc* psf, const float* d_pix, const int64_t* d_offsets,
                             const int32_t* d_ny, const int32_t* d_nx, const float* d_inv_var, const double* d_theta0,
                             uint32_t free_mask, int32_t n_cutouts, const cp_lm_opts* opts, double* d_theta_out,
                             double* d_cov_out, cp_lm_stat* d_stats) {
    if (!ctx) return CP_ESTATE;
    return guarded(ctx, [&] {
        CP_HIP(hipSetDevice(ctx->device));
        lm_batch_device(ctx, psf, d_pix, d_offsets, d_ny, d_nx, d_inv_var, d_theta0, free_mask, n_cutouts, opts,
                        d_theta_out, d_cov_out, d_stats, CP_PREC_MIXED);
    });
}

int cp_fit_star_batch_device_noalloc(cp_ctx* ctx, const cp_psf_desc* psf, const float* d_pix,
                                     const int64_t* d_offsets, const int32_t* d_ny, const int32_t* d_nx,
                                     const float* d_inv_var, const double* d_theta0, uint32_t free_mask,
                                     int32_t n_cutouts, const cp_lm_opts* opts, double* d_theta_out,
                                     double* d_cov_out, cp_lm_stat* d_stats) {
    if (!ctx) return CP_ESTATE;
    return guarded(ctx, [&] {
        CP_ARG(n_cutouts >= 0, "negative cutout count");
        CP_HIP(hipSetDevice(ctx->device));
        lm_batch_noalloc(ctx, psf, d_pix, d_offsets, d_ny, d_nx, d_inv_var, d_theta0, free_mask, n_cutouts, opts,
                         d_theta_out, d_cov_out, d_stats);
    });
}

int cp_fit_star_batch(cp_ctx* ctx, const cp_psf_desc* psf, const float* pix, const int64_t* offsets,
                      const int32_t* ny, const int32_t* nx, const float* inv_var, const double* theta0,
                      uint32_t free_mask, int32_t n, const cp_lm_opts* opts, double* theta_out, double* cov_out,
                      cp_lm_stat* stats) {
    if (!ctx) return CP_ESTATE;
    return guarded(ctx, [&] {
        CP_ARG(n >= 0, "negative cutout count");
        CP_ARG(psf != nullptr, "psf is NULL");
        CP_HIP(hipSetDevice(ctx->device));
        hipStream_t st = ctx->stream;
        if (n == 0) return;
        int64_t total = 0;
        for (int c = 0; c < n; ++c) total = std::max<int64_t>(total, offsets[c] + (int64_t)ny[c] * nx[c]);
        const int npf = lm_nparams(psf->kind);
        int np = 0;
        for (int k = 0; k < npf; ++k) np += (free_mask >> k) & 1u;
        DevBuf dp(ctx, total * 4), dof(ctx, (size_t)n * 8), dny(ctx, (size_t)n * 4), dnx(ctx, (size_t)n * 4),
            dth0(ctx, (size_t)n * npf * 8), dth(ctx, (size_t)n * npf * 8), dst(ctx, (size_t)n * sizeof(cp_lm_stat)),
            div, dcov;
        CP_HIP(hipMemcpyAsync(dp.get(), pix, total * 4, hipMemcpyHostToDevice, st));
        CP_HIP(hipMemcpyAsync(dof.get(), offsets, (size_t)n * 8, hipMemcpyHostToDevice, st));
        CP_HIP(hipMemcpyAsync(dny.get(), ny, (size_t)n * 4, hipMemcpyHostToDevice, st));
        CP_HIP(hipMemcpyAsync(dnx.get(), nx, (size_t)n * 4, hipMemcpyHostToDevice, st));
        CP_HIP(hipMemcpyAsync(dth0.get(), theta0, (size_t)n * npf * 8, hipMemcpyHostToDevice, st));
        if (inv_var) {
            div.alloc(ctx, total * 4);
            CP_HIP(hipMemcpyAsync(div.get(), inv_var, total * 4, hipMemcpyHostToDevice, st));
        }
        if (cov_out) dcov.alloc(ctx, (size_t)n * std::max(np, 1) * std::max(np, 1) * 8);
        lm_batch_device(ctx, psf, dp.as<float>(), dof.as<int64_t>(), dny.as<int32_t>(), dnx.as<int32_t>(),
                        inv_var ? div.as<float>() : nullptr, dth0.as<double>(), free_mask, n, opts, dth.as<double>(),
                        cov_out ? dcov.as<double>() : nullptr, dst.as<cp_lm_stat>(), CP_PREC_MIXED);
        CP_HIP(hipMemcpyAsync(theta_out, dth.get(), (size_t)n * npf * 8, hipMemcpyDeviceToHost, st));
        CP_HIP(hipMemcpyAsync(stats, dst.get(), (size_t)n * sizeof(cp_lm_stat), hipMemcpyDeviceToHost, st));
        if (cov_out)
            CP_HIP(hipMemcpyAsync(cov_out, dcov.get(), (size_t)n * np * np * 8, hipMemcpyDeviceToHost, st));
        CP_HIP(hipStreamSynchronize(st));
    });
}

}  // extern "C"
