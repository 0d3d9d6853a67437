// lm_batch.hip -- batched per-star Levenberg-Marquardt (fit_star over many
// uneven cutouts) on gfx950.
//
// One wavefront owns one cutout for its whole fit: each accum! pass strides the
// wave's lanes over the cutout (read once from HBM, then L1/L2-resident),
// evaluates the model and its gradient per pixel in the storage precision T
// (float32 in the mixed mode) and accumulates the normal equations A = J'WJ,
// b = J'Wr and the cost in float64 with wave-shuffle reductions.  Every lane then runs the same
// lm_irls step (float64 NP x NP Cholesky) so control flow stays uniform.
// Waves pull cutouts from an integer work counter until it passes the end.
//
// References: fit_star src/psf/psf_fitting.jl:229-308 with the specialised
// accumulators :314-451 (CircularGaussianPSF) and :772-862 (CircularMoffatPSF);
// lm_irls src/levenberg_marquardt.jl:431-680; damp! :74-99; covariance! :366-388.
#include <cstring>

#include "sim_kernels.h"
#include "psf_host.h"

namespace cp {

struct LMConst {
    int kind, npf;          // model kind, full parameter count
    int fidx[6];            // free slot -> parameter index (getproperties order)
    double lam_init, lam_up, lam_down, lam_min, lam_max;
    int max_iter, damping, cov_mode;
    double x_tol, f_tol, g_tol, min_diag;
    int have_iv;
};

template <int NP>
struct NE {
    double A[NP * (NP + 1) / 2];
    double b[NP];
    double cost;
};

__device__ __forceinline__ int tri(int i, int j) {  // packed upper triangle, i <= j
    return j * (j + 1) / 2 + i;
}

// Full getproperties-ordered gradient entry k of the model at (ddy, ddx).
template <typename T, int KIND>
__device__ __forceinline__ void model_fg(const T* prm, T ddy, T ddx, T& f, T G[6]) {
    if constexpr (KIND == KIND_GAUSS) {
        // (y, x, fwhm, flux, bkg), evaluate_fg parametric_models.jl:92-113
        T fwhm = prm[2], flux = prm[3], bkg = prm[4];
        T f2 = fwhm * fwhm;
        T sq = (ddx * ddx + ddy * ddy) / f2;
        T norm = -(T(M_PI) * f2 / T(CP_GAUSS_PRE));
        T amp = flux / norm;
        T g = exp(T(CP_GAUSS_PRE) * sq);
        T Ag = amp * g;
        f = fma(amp, g, bkg);
        T gf2 = T(CP_GAUSS_PRE) / f2;
        G[0] = T(-2) * Ag * gf2 * ddy;
        G[1] = T(-2) * Ag * gf2 * ddx;
        G[2] = T(-2) * Ag * (T(1) + T(CP_GAUSS_PRE) * sq) / fwhm;
        G[3] = g / norm;
        G[4] = T(1);
        G[5] = T(0);
    } else {
        // (y, x, alpha, beta, flux, bkg), evaluate_fg parametric_models.jl:704-724
        T a = prm[2], be = prm[3], flux = prm[4], bkg = prm[5];
        T a2 = a * a;
        T r2 = ddx * ddx + ddy * ddy;
        T u = T(1) + r2 / a2;
        T lu = log(u);
        T prof = exp(-be * lu);
        T norm = T(M_PI) * a2 / (be - T(1));
        T amp = flux / norm;
        T Ag = amp * prof;
        f = fma(amp, prof, bkg);
        G[0] = T(2) * Ag * be * ddy / (a2 * u);
        G[1] = T(2) * Ag * be * ddx / (a2 * u);
        G[2] = Ag * (T(-2) / a + T(2) * be * r2 / (a * a2 * u));
        G[3] = Ag * (T(1) / (be - T(1)) - lu);
        G[4] = prof / norm;
        G[5] = T(1);
    }
}

__device__ __forceinline__ double pick6(const double* v, int k) {
    return k == 0 ? v[0] : k == 1 ? v[1] : k == 2 ? v[2] : k == 3 ? v[3] : k == 4 ? v[4] : v[5];
}

// accum!: cost, b = J'W r, A = J'W J over the cutout (lanes stride over its
// pixels; the cutout stays L1/L2-resident across the passes of one fit).
template <typename T, int KIND, int NP>
__device__ __forceinline__ void accum(const LMConst& C, const double* full, const float* __restrict__ pix,
                                      const float* __restrict__ iv, int ny, int npx, NE<NP>& out) {
    const int lane = threadIdx.x & 63;
    T prm[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) prm[k] = (T)full[k];
    NE<NP> acc;
#pragma unroll
    for (int i = 0; i < NP * (NP + 1) / 2; ++i) acc.A[i] = 0;
#pragma unroll
    for (int i = 0; i < NP; ++i) acc.b[i] = 0;
    acc.cost = 0;
    for (int m = lane; m < npx; m += 64) {
        const int col = m / ny;
        const T pi = T(m - col * ny + 1), pj = T(col + 1);
        double w = 1.0;
        if (C.have_iv) {
            float wv = iv[m];
            w = (isfinite(wv) && wv > 0.0f) ? (double)wv : 0.0;
        }
        T f, G[6];
        model_fg<T, KIND>(prm, pi - prm[0], pj - prm[1], f, G);
        const T r = f - (T)pix[m];
        const double wr = w * (double)r;
        acc.cost = fma(wr, (double)r, acc.cost);
        const double G6[6] = {(double)G[0], (double)G[1], (double)G[2], (double)G[3], (double)G[4], (double)G[5]};
        double g[NP];
#pragma unroll
        for (int j = 0; j < NP; ++j) g[j] = pick6(G6, C.fidx[j]);
#pragma unroll
        for (int j = 0; j < NP; ++j) {
            acc.b[j] = fma(wr, g[j], acc.b[j]);
#pragma unroll
            for (int i = 0; i <= j; ++i) acc.A[tri(i, j)] = fma(w * g[i], g[j], acc.A[tri(i, j)]);
        }
    }
#pragma unroll
    for (int i = 0; i < NP * (NP + 1) / 2; ++i) out.A[i] = wave_sum(acc.A[i]);
#pragma unroll
    for (int i = 0; i < NP; ++i) out.b[i] = wave_sum(acc.b[i]);
    out.cost = wave_sum(acc.cost);
}

template <int NP>
__device__ __forceinline__ double Aget(const NE<NP>& e, int i, int j) {
    return i <= j ? e.A[tri(i, j)] : e.A[tri(j, i)];
}

// Cholesky of a packed symmetric NP x NP; false if not positive definite.
template <int NP>
__device__ bool chol(const double* Ap, double L[NP][NP]) {
#pragma unroll
    for (int j = 0; j < NP; ++j) {
        double s = Ap[tri(j, j)];
#pragma unroll
        for (int k = 0; k < j; ++k) s -= L[j][k] * L[j][k];
        if (!(s > 0)) return false;
        double dd = sqrt(s);
        L[j][j] = dd;
#pragma unroll
        for (int i = j + 1; i < NP; ++i) {
            double t = Ap[tri(j, i)];
#pragma unroll
            for (int k = 0; k < j; ++k) t -= L[i][k] * L[j][k];
            L[i][j] = t / dd;
        }
    }
    return true;
}

template <int NP>
__device__ void chol_solve(const double L[NP][NP], const double* b, double* x) {
    double y[NP];
#pragma unroll
    for (int i = 0; i < NP; ++i) {
        double s = b[i];
#pragma unroll
        for (int k = 0; k < i; ++k) s -= L[i][k] * y[k];
        y[i] = s / L[i][i];
    }
#pragma unroll
    for (int i = NP - 1; i >= 0; --i) {
        double s = y[i];
#pragma unroll
        for (int k = i + 1; k < NP; ++k) s -= L[k][i] * x[k];
        x[i] = s / L[i][i];
    }
}

// Symmetric inverse: Cholesky, else Jacobi pseudo-inverse (covariance!).
template <int NP>
__device__ void sym_inv(const NE<NP>& e, double* out) {
    double L[NP][NP];
    double Ap[NP * (NP + 1) / 2];
#pragma unroll
    for (int i = 0; i < NP * (NP + 1) / 2; ++i) Ap[i] = e.A[i];
    if (chol<NP>(Ap, L)) {
#pragma unroll
        for (int j = 0; j < NP; ++j) {
            double ej[NP], col[NP];
#pragma unroll
            for (int i = 0; i < NP; ++i) ej[i] = (i == j);
            chol_solve<NP>(L, ej, col);
#pragma unroll
            for (int i = 0; i < NP; ++i) out[i * NP + j] = col[i];
        }
        return;
    }
    double A[NP][NP], E[NP][NP];
#pragma unroll
    for (int i = 0; i < NP; ++i)
#pragma unroll
        for (int j = 0; j < NP; ++j) {
            A[i][j] = Aget<NP>(e, i, j);
            E[i][j] = (i == j);
        }
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0;
#pragma unroll
        for (int i = 0; i < NP; ++i)
#pragma unroll
            for (int j = i + 1; j < NP; ++j) off += A[i][j] * A[i][j];
        if (off < 1e-300) break;
#pragma unroll
        for (int i = 0; i < NP; ++i)
#pragma unroll
            for (int j = i + 1; j < NP; ++j) {
                double aij = A[i][j];
                if (aij == 0) continue;
                double th = (A[j][j] - A[i][i]) / (2 * aij);
                double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1));
                double c = 1 / sqrt(t * t + 1), s = t * c;
#pragma unroll
                for (int k = 0; k < NP; ++k) {
                    double aki = A[k][i], akj = A[k][j];
                    A[k][i] = c * aki - s * akj;
                    A[k][j] = s * aki + c * akj;
                }
#pragma unroll
                for (int k = 0; k < NP; ++k) {
                    double aik = A[i][k], ajk = A[j][k];
                    A[i][k] = c * aik - s * ajk;
                    A[j][k] = s * aik + c * ajk;
                }
#pragma unroll
                for (int k = 0; k < NP; ++k) {
                    double vki = E[k][i], vkj = E[k][j];
                    E[k][i] = c * vki - s * vkj;
                    E[k][j] = s * vki + c * vkj;
                }
            }
    }
    double smax = 0;
#pragma unroll
    for (int i = 0; i < NP; ++i) smax = fmax(smax, fabs(A[i][i]));
    double tol = DEPS * NP * smax;
#pragma unroll
    for (int i = 0; i < NP; ++i)
#pragma unroll
        for (int j = 0; j < NP; ++j) {
            double s = 0;
#pragma unroll
            for (int k = 0; k < NP; ++k)
                if (fabs(A[k][k]) > tol) s += E[i][k] * E[j][k] / A[k][k];
            out[i * NP + j] = s;
        }
}

template <typename T, int KIND, int NP>
__global__ __launch_bounds__(BLOCK) void k_lm_batch(LMConst C, const float* __restrict__ pix,
                                                    const int64_t* __restrict__ offsets,
                                                    const int32_t* __restrict__ cny, const int32_t* __restrict__ cnx,
                                                    const float* __restrict__ iv, const double* __restrict__ theta0,
                                                    int n, double* __restrict__ theta_out,
                                                    double* __restrict__ cov_out, cp_lm_stat* __restrict__ stats,
                                                    int* __restrict__ work) {
    const int lane = threadIdx.x & 63;
    for (;;) {
        int c = 0;
        if (lane == 0) c = atomicAdd(work, 1);
        c = __shfl(c, 0, 64);
        if (c >= n) return;
        const int ny = cny[c], nx = cnx[c];
        const int npx = ny * nx;
        const int64_t off = offsets[c];
        double full[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int k = 0; k < 6; ++k)
            if (k < C.npf) full[k] = theta0[(size_t)c * C.npf + k];
        cp_lm_stat st{};
        const float* cpix = pix + off;
        const float* civ = C.have_iv ? iv + off : nullptr;
        // input contract (_prepare_fit_star_inputs, psf_fitting.jl:36-101; lm_irls :451-472)
        bool bad = npx - NP <= 0;
        if (C.have_iv && !bad) {
            int nvalid = 0;
            for (int m = lane; m < npx; m += 64) {
                float wv = civ[m];
                nvalid += (isfinite(wv) && wv > 0.0f) ? 1 : 0;
            }
            bad = !(wave_sum_int(nvalid) > NP);
        }
        if (bad) {
            if (lane == 0) {
                st.status = CP_EARG;
                for (int k = 0; k < C.npf; ++k) theta_out[(size_t)c * C.npf + k] = full[k];
                stats[c] = st;
            }
            continue;
        }
        const int dof = npx - NP;
        double x[NP];
#pragma unroll
        for (int j = 0; j < NP; ++j) x[j] = pick6(full, C.fidx[j]);
        auto to_full = [&](const double* xv, double* fo) {
#pragma unroll
            for (int k = 0; k < 6; ++k) fo[k] = full[k];
#pragma unroll
            for (int j = 0; j < NP; ++j) {
                const int fi = C.fidx[j];
#pragma unroll
                for (int k = 0; k < 6; ++k)
                    if (k == fi) fo[k] = xv[j];
            }
        };
        double lam = C.lam_init;
        NE<NP> cur, cand;
        double fx[6];
        to_full(x, fx);
        accum<T, KIND, NP>(C, fx, cpix, civ, ny, npx, cur);
        const double cost_init = cur.cost;
        double D[NP];
#pragma unroll
        for (int i = 0; i < NP; ++i) D[i] = sqrt(fmax(cur.A[tri(i, i)], 1e-6));
        bool xconv = false, fconv = false, gconv = false, conv = false;
        int iter = 0;
        while (iter < C.max_iter) {
            ++iter;
            double gcos = 0;
#pragma unroll
            for (int i = 0; i < NP; ++i) gcos = fmax(gcos, fabs(cur.b[i]) / (sqrt(cur.A[tri(i, i)] * cur.cost) + DEPS));
            if (gcos <= C.g_tol) {
                gconv = conv = true;
                break;
            }
            double Ad[NP * (NP + 1) / 2];
#pragma unroll
            for (int i = 0; i < NP * (NP + 1) / 2; ++i) Ad[i] = cur.A[i];
#pragma unroll
            for (int i = 0; i < NP; ++i) {
                double& dii = Ad[tri(i, i)];
                if (C.damping == CP_DAMP_MARQUARDT) dii += lam * fmax(dii, C.min_diag);
                else if (C.damping == CP_DAMP_LEVENBERG) dii += lam;
            }
            double L[NP][NP];
            if (!chol<NP>(Ad, L)) {
                lam = fmin(lam * C.lam_up, C.lam_max);
                continue;
            }
            double dlt[NP];
            chol_solve<NP>(L, cur.b, dlt);
            double db = 0, dAd = 0;
#pragma unroll
            for (int i = 0; i < NP; ++i) {
                dlt[i] = -dlt[i];
                db += dlt[i] * cur.b[i];
            }
#pragma unroll
            for (int i = 0; i < NP; ++i)
#pragma unroll
                for (int j = 0; j < NP; ++j) dAd += dlt[i] * Aget<NP>(cur, i, j) * dlt[j];
            const double prered = -2 * db - dAd;
            double xc[NP];
#pragma unroll
            for (int i = 0; i < NP; ++i) xc[i] = x[i] + dlt[i];
            to_full(xc, fx);
            accum<T, KIND, NP>(C, fx, cpix, civ, ny, npx, cand);
            const bool accepted = cand.cost < cur.cost;
            const double dcost = cur.cost - cand.cost;
            const double ratio = prered > 0 ? dcost / prered : 0.0;
            fconv = fabs(dcost) <= C.f_tol * (fabs(cur.cost) + C.f_tol) &&
                    prered <= C.f_tol * (fabs(cur.cost) + C.f_tol) && 0.5 * ratio <= 1;
            double nd = 0, nx2 = 0;
#pragma unroll
            for (int i = 0; i < NP; ++i) {
                nd += (D[i] * dlt[i]) * (D[i] * dlt[i]);
                nx2 += (D[i] * x[i]) * (D[i] * x[i]);
            }
            xconv = sqrt(nd) <= C.x_tol * (sqrt(nx2) + C.x_tol) && (accepted || fconv);
            if (accepted) {
                lam = fmax(lam / C.lam_down, C.lam_min);
#pragma unroll
                for (int i = 0; i < NP; ++i) x[i] = xc[i];
                cur = cand;
#pragma unroll
                for (int i = 0; i < NP; ++i) D[i] = fmax(D[i], sqrt(cur.A[tri(i, i)]));
                if (xconv || fconv) {
                    conv = true;
                    break;
                }
            } else {
                if (xconv || fconv) {
                    conv = true;
                    break;
                }
                lam = fmin(lam * C.lam_up, C.lam_max);
            }
        }
        if (lane == 0) {
            double fo[6];
            to_full(x, fo);
            for (int k = 0; k < C.npf; ++k) theta_out[(size_t)c * C.npf + k] = fo[k];
            if (cov_out && C.cov_mode != CP_COV_NONE) {
                double cv[NP * NP];
                sym_inv<NP>(cur, cv);
                double fac = C.cov_mode == CP_COV_REWEIGHTED ? cur.cost / dof : 1.0;
                for (int i = 0; i < NP * NP; ++i) cov_out[(size_t)c * NP * NP + i] = fac * cv[i];
            }
            st.minimum = cur.cost;
            st.cost_init = cost_init;
            st.lambda_final = lam;
            st.chisq = cur.cost / dof;
            st.iterations = iter;
            st.converged = conv;
            st.x_converged = xconv;
            st.f_converged = fconv;
            st.g_converged = gconv;
            st.status = 0;
            stats[c] = st;
        }
    }
}

template <typename T, int KIND>
void launch_lm(int np, dim3 g, hipStream_t s, const LMConst& C, const float* pix, const int64_t* off,
               const int32_t* ny, const int32_t* nx, const float* iv, const double* th0, int n, double* th,
               double* cov, cp_lm_stat* st, int* work) {
#define CP_LM(NPV) launch_n("k_lm_batch", k_lm_batch<T, KIND, NPV>, g, dim3(BLOCK), 0, s, C, pix, off, ny, nx, iv, th0, n, th, cov, st, work)
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

void lm_batch_device(cp_ctx* ctx, const cp_psf_desc* psf, const float* d_pix, const int64_t* d_off,
                     const int32_t* d_ny, const int32_t* d_nx, const float* d_iv, const double* d_th0,
                     uint32_t free_mask, int n, const cp_lm_opts* o, double* d_th, double* d_cov,
                     cp_lm_stat* d_st, int precision) {
    CP_ARG(psf && o, "NULL argument");
    CP_ARG(psf->kind == CP_PSF_CIRCULAR_GAUSSIAN || psf->kind == CP_PSF_CIRCULAR_MOFFAT,
           "batched fit_star supports CircularGaussianPSF and CircularMoffatPSF");
    CP_ARG(o->max_iter >= 0, "max_iter must be non-negative");
    const int npf = psf->kind == CP_PSF_CIRCULAR_GAUSSIAN ? 5 : 6;
    LMConst C{};
    C.kind = psf->kind;
    C.npf = npf;
    int np = 0;
    for (int k = 0; k < npf; ++k)
        if (free_mask & (1u << k)) C.fidx[np++] = k;
    CP_ARG(np > 0, "all model parameters are fixed; nothing to fit");
    CP_ARG((free_mask >> npf) == 0, "free mask names a parameter the model does not have");
    C.lam_init = o->lambda_init;
    C.lam_up = o->lambda_up;
    C.lam_down = o->lambda_down;
    C.lam_min = o->lambda_min;
    C.lam_max = o->lambda_max;
    C.max_iter = o->max_iter;
    C.damping = o->damping;
    C.x_tol = o->x_tol;
    C.f_tol = o->f_tol;
    C.g_tol = o->g_tol;
    C.min_diag = o->min_diagonal;
    C.have_iv = d_iv != nullptr;
    C.cov_mode = o->covariance == CP_COV_DEFAULT ? (d_iv ? CP_COV_KNOWN : CP_COV_REWEIGHTED) : o->covariance;
    if (n == 0) return;
    hipStream_t st = ctx->stream;
    DevBuf work(ctx, 4);
    CP_HIP(hipMemsetAsync(work.get(), 0, 4, st));
    dim3 g((unsigned)std::min<int64_t>((n + NWAVE - 1) / NWAVE, (int64_t)std::max(ctx->n_cu, 1) * 8));
    if (psf->kind == CP_PSF_CIRCULAR_GAUSSIAN) {
        if (precision == CP_PREC_FP64)
            launch_lm<double, KIND_GAUSS>(np, g, st, C, d_pix, d_off, d_ny, d_nx, d_iv, d_th0, n, d_th, d_cov, d_st, work.as<int>());
        else
            launch_lm<float, KIND_GAUSS>(np, g, st, C, d_pix, d_off, d_ny, d_nx, d_iv, d_th0, n, d_th, d_cov, d_st, work.as<int>());
    } else {
        if (precision == CP_PREC_FP64)
            launch_lm<double, KIND_MOFFAT>(np, g, st, C, d_pix, d_off, d_ny, d_nx, d_iv, d_th0, n, d_th, d_cov, d_st, work.as<int>());
        else
            launch_lm<float, KIND_MOFFAT>(np, g, st, C, d_pix, d_off, d_ny, d_nx, d_iv, d_th0, n, d_th, d_cov, d_st, work.as<int>());
    }
    CP_HIP(hipStreamSynchronize(st));
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
        const int npf = psf->kind == CP_PSF_CIRCULAR_GAUSSIAN ? 5 : 6;
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
