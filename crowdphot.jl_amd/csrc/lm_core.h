// lm_core.h -- one wavefront's damped Gauss-Newton fit of one star
// (fit_star + lm_irls) over a pixel source, shared by the batched cutout fit
// and the sequential multi-pass photometry.
//
// The wave's lanes stride over the fit region, evaluate the model and its
// gradient per pixel in the storage precision T (float32 in the mixed mode)
// and accumulate the normal equations A = J'WJ, b = J'Wr and the cost in
// float64 with wave-shuffle reductions.  Every lane then runs the same
// lm_irls step (float64 NP x NP Cholesky), so control flow stays uniform.
//
// References: fit_star src/psf/psf_fitting.jl:229-308 with the specialised
// accumulators :314-451 (CircularGaussianPSF), :772-862 (CircularMoffatPSF),
// :1086-1173 (ImagePSF) and :1177-1210 (GriddedPSFModel); lm_irls
// src/levenberg_marquardt.jl:431-680; damp! :74-99; covariance! :366-388.
#pragma once

#include "sim_kernels.h"

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
        T if2 = T(1) / f2, ifw = T(1) / fwhm;   // loop invariant in accum!
        T sq = (ddx * ddx + ddy * ddy) * if2;
        T norm = -(T(M_PI) * f2 / T(CP_GAUSS_PRE));
        T amp = flux / norm;
        // exp(-4 ln 2 sq) = 2^(-4 sq): float32 takes the hardware exp2
        T g;
        if constexpr (sizeof(T) == 4) g = __builtin_amdgcn_exp2f(T(-4) * sq);
        else g = exp(T(CP_GAUSS_PRE) * sq);
        T Ag = amp * g;
        f = fma(amp, g, bkg);
        T gf2 = T(CP_GAUSS_PRE) / f2;
        G[0] = T(-2) * Ag * gf2 * ddy;
        G[1] = T(-2) * Ag * gf2 * ddx;
        G[2] = T(-2) * Ag * (T(1) + T(CP_GAUSS_PRE) * sq) * ifw;
        G[3] = g / norm;
        G[4] = T(1);
        G[5] = T(0);
    } else {
        // (y, x, alpha, beta, flux, bkg), evaluate_fg parametric_models.jl:704-724
        // The parameter-only factors (1/alpha^2, 1/norm, ...) are loop
        // invariant in accum! and hoisted out of the pixel loop; per pixel
        // one reciprocal of u = 1 + r^2/alpha^2 serves all four gradients.
        T a = prm[2], be = prm[3], flux = prm[4], bkg = prm[5];
        T a2 = a * a;
        T ia = T(1) / a, ia2 = ia * ia;
        T r2 = ddx * ddx + ddy * ddy;
        T u = fma(r2, ia2, T(1));
        // u >= 1: float32 takes the hardware reciprocal, log2 and exp2
        // (u^-beta = 2^(-beta log2 u)), float64 the library functions
        T iu, lu, prof;
        if constexpr (sizeof(T) == 4) {
            iu = __builtin_amdgcn_rcpf(u);
            const T l2 = __builtin_amdgcn_logf(u);
            lu = l2 * T(M_LN2);
            prof = __builtin_amdgcn_exp2f(-be * l2);
        } else {
            iu = T(1) / u;
            lu = log(u);
            prof = exp(-be * lu);
        }
        T inorm = (be - T(1)) / (T(M_PI) * a2);
        T amp = flux * inorm;
        T Ag = amp * prof;
        f = fma(amp, prof, bkg);
        T c = T(2) * Ag * be * ia2 * iu;   // 2 A g beta / (alpha^2 u)
        G[0] = c * ddy;
        G[1] = c * ddx;
        G[2] = Ag * (T(-2) * ia) + c * r2 * ia;
        G[3] = Ag * (T(1) / (be - T(1)) - lu);
        G[4] = prof * inorm;
        G[5] = T(1);
    }
}

// The ePSF grids and node layout of a CP_PSF_GRIDDED_IMAGE model (unused by
// the analytic kinds).  M.bkg is 0: the background is a fitted parameter.
template <typename T>
struct LMModel {
    ModelConst<T> M;
    NodeGrid G;
};

// Bilinear node weights of a multi-node ePSF at the star's current position
// (GriddedPSFModel evaluate_fg, gridded_psf.jl:252-271), once per parameter
// vector; nothing for the other kinds.
template <typename T, int KIND>
__device__ __forceinline__ void model_corners(const LMModel<T>& Md, const double* full, CornerRec<T>& cr) {
    if constexpr (KIND == KIND_EPSF) grid_corners<T>(Md.G, Md.M.n_nodes, full[0], full[1], cr);
}

// f and the getproperties-ordered gradient at offset (ddy, ddx) from the
// centre: the analytic formulas above, or for the ePSF (y, x, flux, bkg)
// f = flux s + bkg with G = (flux ds/dy, flux ds/dx, s, 1) through the same
// bicubic / node-blend device code the simultaneous fit uses.
template <typename T, int KIND>
__device__ __forceinline__ void model_eval(const LMModel<T>& Md, const CornerRec<T>& cr, const T* prm, T ddy,
                                           T ddx, T& f, T G[6]) {
    if constexpr (KIND == KIND_EPSF) {
        StarRec<T> r{};
        r.flux = prm[2];
        r.amp = prm[2];
        T gy, gx, gf;
        // one bicubic pass gives the value too: gf is the unit-flux profile s
        star_grad<T, KIND_EPSF>(Md.M, r, ddy, ddx, Md.M.grids, &cr, gy, gx, gf);
        f = fma(prm[2], gf, prm[3]);
        G[0] = gy;
        G[1] = gx;
        G[2] = gf;
        G[3] = T(1);
        G[4] = G[5] = T(0);
    } else {
        model_fg<T, KIND>(prm, ddy, ddx, f, G);
    }
}

__device__ __forceinline__ double pick6(const double* v, int k) {
    return k == 0 ? v[0] : k == 1 ? v[1] : k == 2 ? v[2] : k == 3 ? v[3] : k == 4 ? v[4] : v[5];
}
template <typename T>
__device__ __forceinline__ T pick6t(const T* v, int k) {
    return k == 0 ? v[0] : k == 1 ? v[1] : k == 2 ? v[2] : k == 3 ? v[3] : k == 4 ? v[4] : v[5];
}

// Pixel sources of accum!: pixel m of the fit region (column-major, rows()
// rows; row and col its position, which the caller steps along with m) gives
// its 1-based image coordinates, its data value and its weight (0 for a
// masked pixel).
// A packed column-major cutout fitted in its own coordinates
// (inds = axes(cutout)).
struct CutoutSrc {
    const float* pix;
    const float* iv;   // NULL: unit weights
    int ny, npx;
    __device__ __forceinline__ int rows() const { return ny; }
    __device__ __forceinline__ void at(int m, int row, int col, double& py, double& px, double& v, double& w) const {
        py = (double)(row + 1);
        px = (double)(col + 1);
        v = (double)pix[m];
        w = 1.0;
        if (iv) {
            const float wv = iv[m];
            w = (isfinite(wv) && wv > 0.0f) ? (double)wv : 0.0;
        }
    }
};

// A packed cutout without inv_var: every weight is the constant 1, so the
// weight products of accum! fold away at compile time.
struct CutoutSrcUnit {
    const float* pix;
    int ny, npx;
    __device__ __forceinline__ int rows() const { return ny; }
    __device__ __forceinline__ void at(int m, int row, int col, double& py, double& px, double& v, double& w) const {
        py = (double)(row + 1);
        px = (double)(col + 1);
        v = (double)pix[m];
        w = 1.0;
    }
};

// Row and column of pixel m = lane + 64 k of a column-major region of ny
// rows, stepped with k: one division per pass instead of one per pixel.
struct PixCursor {
    int row, col, dr, dc, ny;
    __device__ __forceinline__ PixCursor(int lane, int rows) : ny(rows > 0 ? rows : 1) {
        col = lane / ny;
        row = lane - col * ny;
        dc = 64 / ny;
        dr = 64 - dc * ny;
    }
    __device__ __forceinline__ void next() {
        row += dr;
        col += dc;
        if (row >= ny) {
            row -= ny;
            ++col;
        }
    }
};

// accum!: cost, b = J'W r, A = J'W J over the fit region (lanes stride over
// its pixels).  The offsets from the star centre are formed in float64 and
// then rounded, so absolute image coordinates keep their precision in the
// float32 mode.
template <typename T, int KIND, int NP, typename Src>
__device__ __forceinline__ void accum(const LMConst& C, const LMModel<T>& Md, const double* full, const Src& S,
                                      NE<NP>& out) {
    const int lane = threadIdx.x & 63;
    T prm[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) prm[k] = (T)full[k];
    CornerRec<T> cr;
    model_corners<T, KIND>(Md, full, cr);
    NE<NP> acc;
#pragma unroll
    for (int i = 0; i < NP * (NP + 1) / 2; ++i) acc.A[i] = 0;
#pragma unroll
    for (int i = 0; i < NP; ++i) acc.b[i] = 0;
    acc.cost = 0;
    PixCursor pc(lane, S.rows());
    for (int m = lane; m < S.npx; m += 64, pc.next()) {
        double py, px, v, w;
        S.at(m, pc.row, pc.col, py, px, v, w);
        T f, G[6];
        model_eval<T, KIND>(Md, cr, prm, (T)(py - full[0]), (T)(px - full[1]), f, G);
        const T r = f - (T)v;
        const double wr = w * (double)r;
        acc.cost = fma(wr, (double)r, acc.cost);
        // the free gradients, selected in the storage precision, then widened
        double g[NP], wg[NP];
#pragma unroll
        for (int j = 0; j < NP; ++j) {
            g[j] = (double)pick6t<T>(G, C.fidx[j]);
            wg[j] = w * g[j];
        }
#pragma unroll
        for (int j = 0; j < NP; ++j) {
            acc.b[j] = fma(wr, g[j], acc.b[j]);
#pragma unroll
            for (int i = 0; i <= j; ++i) acc.A[tri(i, j)] = fma(wg[i], g[j], acc.A[tri(i, j)]);
        }
    }
    // the NA + NP + 1 sums in one reduce-scatter (wave_sum's sums exactly)
    constexpr int NA = NP * (NP + 1) / 2;
    double s[NA + NP + 1];
#pragma unroll
    for (int i = 0; i < NA; ++i) s[i] = acc.A[i];
#pragma unroll
    for (int i = 0; i < NP; ++i) s[NA + i] = acc.b[i];
    s[NA + NP] = acc.cost;
    wave_sum_many<NA + NP + 1>(s);
#pragma unroll
    for (int i = 0; i < NA; ++i) out.A[i] = s[i];
#pragma unroll
    for (int i = 0; i < NP; ++i) out.b[i] = s[NA + i];
    out.cost = s[NA + NP];
}

template <int NP>
__device__ __forceinline__ double Aget(const NE<NP>& e, int i, int j) {
    return i <= j ? e.A[tri(i, j)] : e.A[tri(j, i)];
}

// Cholesky of a packed symmetric NP x NP; false if not positive definite.
// The diagonal of L holds the reciprocals of the pivots (one division per
// column; chol_solve multiplies by them).
template <int NP>
__device__ bool chol(const double* Ap, double L[NP][NP]) {
#pragma unroll
    for (int j = 0; j < NP; ++j) {
        double s = Ap[tri(j, j)];
#pragma unroll
        for (int k = 0; k < j; ++k) s -= L[j][k] * L[j][k];
        if (!(s > 0)) return false;
        const double idd = 1.0 / sqrt(s);
        L[j][j] = idd;
#pragma unroll
        for (int i = j + 1; i < NP; ++i) {
            double t = Ap[tri(j, i)];
#pragma unroll
            for (int k = 0; k < j; ++k) t -= L[i][k] * L[j][k];
            L[i][j] = t * idd;
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
        y[i] = s * L[i][i];
    }
#pragma unroll
    for (int i = NP - 1; i >= 0; --i) {
        double s = y[i];
#pragma unroll
        for (int k = i + 1; k < NP; ++k) s -= L[k][i] * x[k];
        x[i] = s * L[i][i];
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


// Result of one lm_irls run.
template <int NP>
struct LMOut {
    double full[6];   // parameters at the minimiser (getproperties order)
    NE<NP> cur;       // normal equations and cost there
    cp_lm_stat st;    // st.status = CP_EARG: rejected input, full = the start
};

// _prepare_fit_star_inputs' checks (psf_fitting.jl:36-101; lm_irls :451-472)
// and the lm_irls iteration from the full parameter vector full0 over the
// pixel source S.  Wave-uniform: every lane returns the same R.
template <typename T, int KIND, int NP, typename Src>
__device__ void lm_fit(const LMConst& C, const LMModel<T>& Md, const double* full0, const Src& S, LMOut<NP>& R) {
    const int lane = threadIdx.x & 63;
    double full[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < 6; ++k)
        if (k < C.npf) full[k] = full0[k];
    R.st = cp_lm_stat{};
    const int npx = S.npx;
    bool bad = npx - NP <= 0;
    if (C.have_iv && !bad) {
        int nvalid = 0;
        PixCursor pc(lane, S.rows());
        for (int m = lane; m < npx; m += 64, pc.next()) {
            double py, px, v, w;
            S.at(m, pc.row, pc.col, py, px, v, w);
            nvalid += w > 0 ? 1 : 0;
        }
        bad = !(wave_sum_int(nvalid) > NP);
    }
    if (bad) {
        R.st.status = CP_EARG;
#pragma unroll
        for (int k = 0; k < 6; ++k) R.full[k] = full[k];
        return;
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
    NE<NP> cand;
    NE<NP>& cur = R.cur;
    double fx[6];
    to_full(x, fx);
    accum<T, KIND, NP>(C, Md, fx, S, cur);
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
        accum<T, KIND, NP>(C, Md, fx, S, cand);
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
    to_full(x, R.full);
    R.st.minimum = cur.cost;
    R.st.cost_init = cost_init;
    R.st.lambda_final = lam;
    R.st.chisq = cur.cost / dof;
    R.st.iterations = iter;
    R.st.converged = conv;
    R.st.x_converged = xconv;
    R.st.f_converged = fconv;
    R.st.g_converged = gconv;
    R.st.status = 0;
}

// covariance! (levenberg_marquardt.jl:366-388) of the free parameters at the
// minimiser into cv[NP * NP]: inverse of J'WJ, scaled by cost / dof for the
// reweighted estimator.
template <int NP>
__device__ void lm_cov(const LMConst& C, const LMOut<NP>& R, int npx, double* cv) {
    sym_inv<NP>(R.cur, cv);
    const double fac = C.cov_mode == CP_COV_REWEIGHTED ? R.cur.cost / (npx - NP) : 1.0;
#pragma unroll
    for (int i = 0; i < NP * NP; ++i) cv[i] *= fac;
}

inline int lm_nparams(int kind) {
    return kind == CP_PSF_CIRCULAR_GAUSSIAN ? 5 : kind == CP_PSF_CIRCULAR_MOFFAT ? 6 : 4;
}

// LMConst from the options; false (with the reason in *why) for a model the
// kernel does not cover or a bad free mask.
inline bool lm_const(const cp_psf_desc* psf, uint32_t free_mask, const cp_lm_opts* o, bool have_iv, LMConst& C,
                     int& np, const char** why) {
    if (!(psf->kind == CP_PSF_CIRCULAR_GAUSSIAN || psf->kind == CP_PSF_CIRCULAR_MOFFAT ||
          psf->kind == CP_PSF_GRIDDED_IMAGE)) {
        *why = "unknown PSF kind";
        return false;
    }
    if (o->max_iter < 0) {
        *why = "max_iter must be non-negative";
        return false;
    }
    const int npf = lm_nparams(psf->kind);
    C = LMConst{};
    C.kind = psf->kind;
    C.npf = npf;
    np = 0;
    for (int k = 0; k < npf; ++k)
        if (free_mask & (1u << k)) C.fidx[np++] = k;
    if (np == 0) {
        *why = "all model parameters are fixed; nothing to fit";
        return false;
    }
    if ((free_mask >> npf) != 0) {
        *why = "free mask names a parameter the model does not have";
        return false;
    }
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
    C.have_iv = have_iv;
    C.cov_mode = o->covariance == CP_COV_DEFAULT ? (have_iv ? CP_COV_KNOWN : CP_COV_REWEIGHTED) : o->covariance;
    return true;
}

}  // namespace cp
