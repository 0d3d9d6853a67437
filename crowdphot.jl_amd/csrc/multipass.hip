// multipass.hip -- fit_all_stars (src/photometry/psf_photometry_single.jl:267-469)
// on gfx950: sequential multi-pass PSF photometry, every star fitted with
// fit_star on the progressively subtracted residual image.
//
// The reference walks the stars one at a time, brightest first.  A star's fit
// reads and writes the residual only inside its +-fit_rad box, so stars whose
// boxes do not overlap commute.  Each pass therefore splits the flux order
// into levels -- level(j) = 1 + the highest level of any earlier star whose
// box overlaps j's box -- and fits all stars of one level at once, one
// wavefront per star, the levels in order.  That reproduces the sequential
// result: each star sees exactly the residual its predecessors in flux order
// left over its box, and no other star of its level touches that box.
//
// Per star the wave stages its box of the float64 residual in LDS, adds the
// star's previous model back (passes after the first), runs lm_fit on the LDS
// copy in absolute image coordinates, subtracts the fitted model, writes the
// box back and, on the last pass, computes the star's diagnostics
// (_star_diagnostics!, psf_photometry_diagnostics.jl:154-231) over that box.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "diag_host.h"
#include "lm_core.h"
#include "psf_host.h"

namespace cp {
namespace {

// Largest per-star box the LDS staging takes: (2 ceil(fit_rad) + 2)^2 pixels
// of 20 bytes (residual, model, inverse variance) within 64 KiB.
constexpr int MP_LDS_MAX = 64 * 1024;
constexpr int MP_BYTES_PER_PIXEL = 8 + 8 + 4;

struct MPArgs {
    LMConst C;
    uint32_t free_mask;
    double fixv[6];          // fixed parameter values (NaN: the star's own)
    double* R;               // residual image, float64, column-major ny x nx
    const float* image;      // the input image (diagnostics)
    const float* iv;         // inverse variance or NULL
    int ny, n, npf;
    double* prm;             // n x npf parameters (getproperties order)
    double* err;             // n x npf 1-sigma errors
    const int4* box;         // per star: ylo, yhi, xlo, xhi (1-based, inside the image)
    uint8_t* valid;
    uint8_t* conv;
    double* chisq;
    int* n_iter;
    int* n_failed;
    int first, last;         // pass flags
    const double* K;         // spread_model disk (2 kh + 1)^2, or NULL
    int kh;
    double* diag;            // 6 x n: qfit, qfit_expected, qfit_z, crowding, spread_model(_err)
    int maxb;                // LDS box capacity in pixels
};

// A star's box of the residual staged in LDS: pixel m is image pixel
// (ylo + m % by, xlo + m / by); weights from the staged inverse variance.
struct BoxSrc {
    const double* cut;
    const float* ivs;
    int by, ylo, xlo, npx;
    __device__ __forceinline__ void at(int m, double& py, double& px, double& v, double& w) const {
        const int col = m / by;
        py = (double)(ylo + m - col * by);
        px = (double)(xlo + col);
        v = cut[m];
        w = 1.0;
        if (ivs) {
            const float wv = ivs[m];
            w = (isfinite(wv) && wv > 0.0f) ? (double)wv : 0.0;
        }
    }
};

template <typename T, int KIND>
__device__ __forceinline__ double model_value(const LMModel<T>& Md, const CornerRec<T>& cr, const T* p,
                                              const double* full, double py, double px) {
    T f, G[6];
    model_eval<T, KIND>(Md, cr, p, (T)(py - full[0]), (T)(px - full[1]), f, G);
    return (double)f;
}

template <typename T, int KIND, int NP>
__global__ __launch_bounds__(64) void k_mp_level(MPArgs A, LMModel<T> Md, const int* __restrict__ list, int cnt) {
    extern __shared__ double lds[];
    if ((int)blockIdx.x >= cnt) return;
    const int lane = threadIdx.x;
    const int s = list[blockIdx.x];
    const int4 bb = A.box[s];
    const int ylo = bb.x, xlo = bb.z, by = bb.y - bb.x + 1;
    const int nb = by * (bb.w - bb.z + 1);
    double* cut = lds;
    double* ms = lds + A.maxb;
    float* ivs = (float*)(lds + 2 * A.maxb);
    const int npf = A.npf;
    const size_t NY = (size_t)A.ny;
    double cur[6] = {0, 0, 0, 0, 0, 0}, full0[6];
    T pc[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        if (k < npf) cur[k] = A.prm[(size_t)s * npf + k];
        const bool fixed_here = !((A.free_mask >> k) & 1u) && !isnan(A.fixv[k]);
        full0[k] = fixed_here ? A.fixv[k] : cur[k];
        pc[k] = (T)cur[k];
    }
    CornerRec<T> cc;
    model_corners<T, KIND>(Md, cur, cc);
    // stage the box; passes after the first add the star's model back (add_star!)
    for (int m = lane; m < nb; m += 64) {
        const int col = m / by, gy = ylo + m - col * by, gx = xlo + col;
        const size_t q = (size_t)(gy - 1) + (size_t)(gx - 1) * NY;
        double v = A.R[q];
        if (!A.first) v += model_value<T, KIND>(Md, cc, pc, cur, gy, gx);
        cut[m] = v;
        if (A.iv) ivs[m] = A.iv[q];
    }
    __syncthreads();
    const BoxSrc S{cut, A.iv ? ivs : nullptr, by, ylo, xlo, nb};
    LMOut<NP> R;
    lm_fit<T, KIND, NP>(A.C, Md, full0, S, R);
    if (R.st.status != 0) {
        // the reference's catch: undo the add-back, mark the star invalid
        if (!A.first)
            for (int m = lane; m < nb; m += 64) {
                const int col = m / by, gy = ylo + m - col * by, gx = xlo + col;
                A.R[(size_t)(gy - 1) + (size_t)(gx - 1) * NY] = cut[m] - model_value<T, KIND>(Md, cc, pc, cur, gy, gx);
            }
        if (lane == 0) {
            A.valid[s] = 0;
            atomicAdd(A.n_failed, 1);
        }
        return;
    }
    // subtract the fitted model (subtract_star! over the same box)
    T pn[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) pn[k] = (T)R.full[k];
    CornerRec<T> cn;
    model_corners<T, KIND>(Md, R.full, cn);
    for (int m = lane; m < nb; m += 64) {
        const int col = m / by, gy = ylo + m - col * by, gx = xlo + col;
        const double f = model_value<T, KIND>(Md, cn, pn, R.full, gy, gx);
        const double r = cut[m] - f;
        A.R[(size_t)(gy - 1) + (size_t)(gx - 1) * NY] = r;
        cut[m] = r;
        ms[m] = f;
    }
    if (lane == 0) {
        for (int k = 0; k < npf; ++k) A.prm[(size_t)s * npf + k] = R.full[k];
        if (A.C.cov_mode != CP_COV_NONE) {   // _extract_errors! (psf_photometry_single.jl:210-222)
            double cv[NP * NP];
            lm_cov<NP>(A.C, R, nb, cv);
            for (int j = 0; j < NP; ++j) A.err[(size_t)s * npf + A.C.fidx[j]] = sqrt(fmax(0.0, cv[j * NP + j]));
        }
        A.conv[s] = R.st.converged;
        A.chisq[s] = R.st.chisq;
        A.n_iter[s] += R.st.iterations;
    }
    if (!A.last) return;
    __syncthreads();
    // _star_diagnostics! on this star's residual (neighbours and its own
    // model removed), the input image and its model over the box
    const int rf = npf - 2, rb = npf - 1;
    const double flux = R.full[rf], bkg = R.full[rb];
    if (!(flux > 0)) return;
    const double inv_flux = 1.0 / flux;
    const int kw = 2 * A.kh + 1, bx = nb / by;
    double qfit = 0, nclean = 0, ndirty = 0, dcrowd = 0, Sp = 0, Sq = 0, Sp0 = 0, Sq0 = 0, SqG = 0;
    double ssum = 0, s2sum = 0, ngood = 0;
    for (int m = lane; m < nb; m += 64) {
        const int ix = m / by, iy = m - ix * by;
        const double wp = A.iv ? (double)ivs[m] : 1.0;
        if (!(isfinite(wp) && wp > 0)) continue;
        const size_t q = (size_t)(ylo + iy - 1) + (size_t)(xlo + ix - 1) * NY;
        const double rq = cut[m], mv = ms[m];
        qfit += fabs(rq);
        const double Pp = (mv - bkg) * inv_flux, wP = wp * Pp, pi = rq + mv - bkg;
        nclean += wP * pi;
        ndirty += wP * ((double)A.image[q] - bkg);
        dcrowd += wP * Pp;
        if (A.K) {   // correlate!(g, ms, K, :zero) (correlation.jl:122-140) at this pixel
            double gm = 0;
            for (int b = -A.kh; b <= A.kh; ++b) {
                const int jx = ix + b;
                if (jx < 0 || jx >= bx) continue;
                for (int a = -A.kh; a <= A.kh; ++a) {
                    const int jy = iy + a;
                    if (jy < 0 || jy >= by) continue;
                    gm += ms[jy + jx * by] * A.K[(a + A.kh) + (b + A.kh) * kw];
                }
            }
            const double gi = (gm - bkg) * inv_flux;
            Sp += wP * pi;
            Sq += wp * gi * pi;
            Sp0 += wP * Pp;
            Sq0 += wp * Pp * gi;
            SqG += wp * gi * gi;
        }
        if (A.iv) {
            const double sg = 1.0 / sqrt(wp);
            ssum += sg;
            s2sum += sg * sg;
            ngood += 1;
        }
    }
    qfit = wave_sum(qfit);
    nclean = wave_sum(nclean);
    ndirty = wave_sum(ndirty);
    dcrowd = wave_sum(dcrowd);
    Sp = wave_sum(Sp);
    Sq = wave_sum(Sq);
    Sp0 = wave_sum(Sp0);
    Sq0 = wave_sum(Sq0);
    SqG = wave_sum(SqG);
    ssum = wave_sum(ssum);
    s2sum = wave_sum(s2sum);
    ngood = wave_sum(ngood);
    if (lane != 0) return;
    const size_t N = (size_t)A.n;
    double* o = A.diag + s;
    o[0] = qfit * inv_flux;
    if (dcrowd > 0 && nclean > 0 && ndirty > 0) o[3 * N] = 2.5 * log10(ndirty / nclean);
    if (A.K && Sp > 0) {
        o[4 * N] = Sq / Sp - Sq0 / Sp0;
        if (A.iv) {
            const double v = Sp * Sp * SqG + Sq * Sq * Sp0 - 2 * Sq * Sp * Sq0;
            o[5 * N] = sqrt(fmax(0.0, v)) / (Sp * Sp);
        }
    }
    if (A.iv) {
        const double c2pi = sqrt(2.0 / M_PI);
        o[N] = c2pi * ssum * inv_flux;
        const int p = NP;
        if (s2sum > 0 && ngood > p) {
            const double dof = sqrt(1.0 - p / ngood);
            o[2 * N] = (qfit - c2pi * dof * ssum) / (sqrt((1 - 2 / M_PI) * s2sum) * dof);
        }
    }
}

// Between-pass validation (psf_photometry_single.jl:427-434): keep converged
// stars with finite position and flux and positive flux.
static __global__ void k_mp_validate(int n, int npf, const double* __restrict__ prm,
                                     const uint8_t* __restrict__ conv, uint8_t* __restrict__ valid) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const double* m = prm + (size_t)i * npf;
    const double f = m[npf - 2];
    valid[i] = valid[i] && conv[i] && isfinite(m[0]) && isfinite(m[1]) && isfinite(f) && f > 0;
}

static __global__ void k_f32_to_f64(int64_t n, const float* __restrict__ a, double* __restrict__ b) {
    for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK)
        b[i] = (double)a[i];
}

static __global__ void k_f64_to_f32(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
    for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK)
        b[i] = (float)a[i];
}

template <typename T, int KIND>
void launch_level(int np, size_t lds, hipStream_t s, const MPArgs& A, const LMModel<T>& Md, const int* list,
                  int cnt) {
#define CP_MP(NPV) launch_n("k_mp_level", k_mp_level<T, KIND, NPV>, dim3(cnt), dim3(64), lds, s, A, Md, list, cnt)
    switch (np) {
        case 1: CP_MP(1); break;
        case 2: CP_MP(2); break;
        case 3: CP_MP(3); break;
        case 4: CP_MP(4); break;
        case 5: CP_MP(5); break;
        default: CP_MP(6); break;
    }
#undef CP_MP
}

// Box of a star centred at (y, x): floor(y - r):ceil(y + r) clamped to the
// image (psf_photometry_single.jl:353-357, _clamp_inds).
struct HostBox {
    long ylo, yhi, xlo, xhi;
};
inline HostBox star_box(double y, double x, double r, int ny, int nx) {
    return {std::max((long)std::floor(y - r), 1L), std::min((long)std::ceil(y + r), (long)ny),
            std::max((long)std::floor(x - r), 1L), std::min((long)std::ceil(x + r), (long)nx)};
}

void fit_all_stars_impl(cp_ctx* ctx, const cp_psf_desc* psf, const float* image, const float* inv_var, int ny,
                        int nx, const double* theta0, int n, double fit_rad, int n_passes, uint32_t free_mask,
                        const double* fixed_values, const cp_lm_opts* o, double spread_model_fwhm, int precision,
                        cp_sim_result* out) {
    CP_ARG(psf && o && out && image, "NULL argument");
    CP_ARG(ny > 0 && nx > 0, "image must be non-empty");
    CP_ARG(n >= 0 && (theta0 || n == 0), "bad star catalogue");
    CP_ARG(n_passes > 0, "n_passes must be positive, got " + std::to_string(n_passes));
    CP_ARG(fit_rad > 0 && std::isfinite(fit_rad), "fit_rad must be positive");
    CP_ARG(precision == CP_PREC_MIXED || precision == CP_PREC_FP64, "unknown precision");
    LMConst C;
    int np = 0;
    const char* why = "";
    CP_ARG(lm_const(psf, free_mask, o, inv_var != nullptr, C, np, &why), why);
    const int npf = C.npf, rf = npf - 2, rb = npf - 1;
    const long W = 2 * (long)std::ceil(fit_rad) + 2;
    CP_ARG(W * W * MP_BYTES_PER_PIXEL <= MP_LDS_MAX, "fit_rad too large for the per-star box (at most " +
                                                         std::to_string((int)std::sqrt(MP_LDS_MAX / MP_BYTES_PER_PIXEL) / 2 - 1) + ")");
    out->n_failed = 0;
    if (n == 0) {   // the reference's empty result (:285-288)
        out->n_passes = 0;
        return;
    }
    const int maxb = (int)(W * W);
    const size_t lds = (size_t)maxb * MP_BYTES_PER_PIXEL;
    hipStream_t st = ctx->stream;
    const size_t NP_ = (size_t)ny * nx;

    // host state mirrored from the device between passes
    std::vector<double> prm(theta0, theta0 + (size_t)n * npf), err((size_t)n * npf);
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < npf; ++k) err[(size_t)i * npf + k] = ((free_mask >> k) & 1u) ? NAN : 0.0;
    std::vector<uint8_t> valid(n, 1);
    std::vector<double> diag0((size_t)6 * n, NAN);

    DevBuf dR(ctx, NP_ * 8), dimg(ctx, NP_ * 4), div, dprm(ctx, prm.size() * 8), derr(ctx, err.size() * 8),
        dbox(ctx, (size_t)n * sizeof(int4)), dvalid(ctx, n), dconv(ctx, n), dchi(ctx, (size_t)n * 8),
        dnit(ctx, (size_t)n * 4), dfail(ctx, 4), ddiag(ctx, (size_t)6 * n * 8), dlist(ctx, (size_t)n * 4), dK;
    CP_HIP(hipMemcpyAsync(dimg.get(), image, NP_ * 4, hipMemcpyHostToDevice, st));
    if (inv_var) {
        div.alloc(ctx, NP_ * 4);
        CP_HIP(hipMemcpyAsync(div.get(), inv_var, NP_ * 4, hipMemcpyHostToDevice, st));
    }
    launch(k_f32_to_f64, dim3((unsigned)std::min<size_t>(cdiv((int64_t)NP_, BLOCK), 65536)), dim3(BLOCK), 0, st,
           (int64_t)NP_, (const float*)dimg.as<float>(), dR.as<double>());
    CP_HIP(hipMemcpyAsync(dprm.get(), prm.data(), prm.size() * 8, hipMemcpyHostToDevice, st));
    CP_HIP(hipMemcpyAsync(derr.get(), err.data(), err.size() * 8, hipMemcpyHostToDevice, st));
    CP_HIP(hipMemcpyAsync(ddiag.get(), diag0.data(), diag0.size() * 8, hipMemcpyHostToDevice, st));
    CP_HIP(hipMemsetAsync(dconv.get(), 0, n, st));
    CP_HIP(hipMemsetAsync(dchi.get(), 0, (size_t)n * 8, st));
    CP_HIP(hipMemsetAsync(dnit.get(), 0, (size_t)n * 4, st));
    CP_HIP(hipMemsetAsync(dfail.get(), 0, 4, st));

    // spread_model reference disk, sized from star 1 (:320-323)
    const int kh = (int)std::ceil(fit_rad);
    const double sfw = spread_model_fwhm > 0 ? spread_model_fwhm : spread_fwhm_host(psf, theta0[0], theta0[1]);
    std::vector<double> hK;
    if (std::isfinite(sfw) && sfw > 0) {
        hK.resize((size_t)(2 * kh + 1) * (2 * kh + 1));
        spread_kernel_host(sfw, kh, hK.data());
        dK.alloc(ctx, hK.size() * 8);
        CP_HIP(hipMemcpyAsync(dK.get(), hK.data(), hK.size() * 8, hipMemcpyHostToDevice, st));
    }

    const HostPsf HP = make_host_psf(psf);
    DevPsf DP;
    upload_psf(ctx, psf, HP, DP);
    const LMModel<double> Md64{model_const<double>(HP, DP, 0.0), DP.G};
    const LMModel<float> Md32{model_const<float>(HP, DP, 0.0), DP.G};

    MPArgs A{};
    A.C = C;
    A.free_mask = free_mask;
    for (int k = 0; k < 6; ++k) A.fixv[k] = (fixed_values && k < npf) ? fixed_values[k] : NAN;
    A.R = dR.as<double>();
    A.image = dimg.as<float>();
    A.iv = inv_var ? div.as<float>() : nullptr;
    A.ny = ny;
    A.n = n;
    A.npf = npf;
    A.prm = dprm.as<double>();
    A.err = derr.as<double>();
    A.box = dbox.as<int4>();
    A.valid = dvalid.as<uint8_t>();
    A.conv = dconv.as<uint8_t>();
    A.chisq = dchi.as<double>();
    A.n_iter = dnit.as<int>();
    A.n_failed = dfail.as<int>();
    A.K = hK.empty() ? nullptr : dK.as<double>();
    A.kh = kh;
    A.diag = ddiag.as<double>();
    A.maxb = maxb;

    std::vector<int> order(n), level(n), lvmap(NP_, 0), cnt, list(n);
    std::vector<int4> box(n);
    for (int pass = 1; pass <= n_passes; ++pass) {
        if (pass > 1) {
            CP_HIP(hipMemcpyAsync(prm.data(), dprm.get(), prm.size() * 8, hipMemcpyDeviceToHost, st));
            CP_HIP(hipMemcpyAsync(valid.data(), dvalid.get(), n, hipMemcpyDeviceToHost, st));
            CP_HIP(hipStreamSynchronize(st));
        }
        // flux order, brightest first; sortperm(rev = true) puts NaN first
        std::iota(order.begin(), order.end(), 0);
        auto key = [&](int i) {
            const double f = prm[(size_t)i * npf + rf];
            return std::isnan(f) ? INFINITY : f;
        };
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return key(a) > key(b); });
        // levels over a per-pixel map of the highest level that covers it
        int nlev = 0;
        for (int idx : order) {
            level[idx] = 0;
            if (!valid[idx]) continue;
            const double y = prm[(size_t)idx * npf], x = prm[(size_t)idx * npf + 1];
            CP_ARG(std::fabs(y) < 1e9 && std::fabs(x) < 1e9,
                   "star " + std::to_string(idx + 1) + " has a non-finite or out-of-range position");
            const HostBox b = star_box(y, x, fit_rad, ny, nx);
            const long by = b.yhi - b.ylo + 1, bx = b.xhi - b.xlo + 1;
            if (by <= 0 || bx <= 0 || by * bx < 3) {   // too few pixels (:358)
                valid[idx] = 0;
                continue;
            }
            int L = 0;
            for (long gx = b.xlo; gx <= b.xhi; ++gx)
                for (long gy = b.ylo; gy <= b.yhi; ++gy) L = std::max(L, lvmap[(size_t)(gy - 1) + (size_t)(gx - 1) * ny]);
            ++L;
            for (long gx = b.xlo; gx <= b.xhi; ++gx)
                for (long gy = b.ylo; gy <= b.yhi; ++gy) lvmap[(size_t)(gy - 1) + (size_t)(gx - 1) * ny] = L;
            level[idx] = L;
            box[idx] = make_int4((int)b.ylo, (int)b.yhi, (int)b.xlo, (int)b.xhi);
            nlev = std::max(nlev, L);
        }
        for (int idx = 0; idx < n; ++idx)   // clear the map for the next pass
            if (level[idx]) {
                const int4 b = box[idx];
                for (int gx = b.z; gx <= b.w; ++gx)
                    for (int gy = b.x; gy <= b.y; ++gy) lvmap[(size_t)(gy - 1) + (size_t)(gx - 1) * ny] = 0;
            }
        cnt.assign(nlev + 2, 0);
        for (int idx = 0; idx < n; ++idx) cnt[level[idx]]++;
        std::vector<int> off(nlev + 2, 0);
        for (int l = 1; l <= nlev; ++l) off[l + 1] = off[l] + cnt[l];
        {
            std::vector<int> fill(off);
            for (int idx = 0; idx < n; ++idx)
                if (level[idx]) list[fill[level[idx]]++] = idx;
        }
        const int total = nlev ? off[nlev + 1] : 0;
        CP_HIP(hipMemcpyAsync(dbox.get(), box.data(), (size_t)n * sizeof(int4), hipMemcpyHostToDevice, st));
        CP_HIP(hipMemcpyAsync(dvalid.get(), valid.data(), n, hipMemcpyHostToDevice, st));
        if (total) CP_HIP(hipMemcpyAsync(dlist.get(), list.data(), (size_t)total * 4, hipMemcpyHostToDevice, st));
        A.first = pass == 1;
        A.last = pass == n_passes;
        for (int l = 1; l <= nlev; ++l) {
            const int* lst = dlist.as<int>() + off[l];
            const bool f64 = precision == CP_PREC_FP64;
            if (psf->kind == CP_PSF_CIRCULAR_GAUSSIAN) {
                if (f64) launch_level<double, KIND_GAUSS>(np, lds, st, A, Md64, lst, cnt[l]);
                else launch_level<float, KIND_GAUSS>(np, lds, st, A, Md32, lst, cnt[l]);
            } else if (psf->kind == CP_PSF_CIRCULAR_MOFFAT) {
                if (f64) launch_level<double, KIND_MOFFAT>(np, lds, st, A, Md64, lst, cnt[l]);
                else launch_level<float, KIND_MOFFAT>(np, lds, st, A, Md32, lst, cnt[l]);
            } else {
                if (f64) launch_level<double, KIND_EPSF>(np, lds, st, A, Md64, lst, cnt[l]);
                else launch_level<float, KIND_EPSF>(np, lds, st, A, Md32, lst, cnt[l]);
            }
        }
        launch(k_mp_validate, dim3(cdiv(n, BLOCK)), dim3(BLOCK), 0, st, n, npf, (const double*)A.prm,
               (const uint8_t*)A.conv, A.valid);
        // the host copies above are pageable and reused next pass
        CP_HIP(hipStreamSynchronize(st));
    }

    // results
    std::vector<uint8_t> conv(n);
    std::vector<double> chisq(n), diag((size_t)6 * n);
    std::vector<int> nit(n);
    int nfail = 0;
    CP_HIP(hipMemcpyAsync(prm.data(), dprm.get(), prm.size() * 8, hipMemcpyDeviceToHost, st));
    CP_HIP(hipMemcpyAsync(err.data(), derr.get(), err.size() * 8, hipMemcpyDeviceToHost, st));
    CP_HIP(hipMemcpyAsync(valid.data(), dvalid.get(), n, hipMemcpyDeviceToHost, st));
    CP_HIP(hipMemcpyAsync(conv.data(), dconv.get(), n, hipMemcpyDeviceToHost, st));
    CP_HIP(hipMemcpyAsync(chisq.data(), dchi.get(), (size_t)n * 8, hipMemcpyDeviceToHost, st));
    CP_HIP(hipMemcpyAsync(nit.data(), dnit.get(), (size_t)n * 4, hipMemcpyDeviceToHost, st));
    CP_HIP(hipMemcpyAsync(diag.data(), ddiag.get(), diag.size() * 8, hipMemcpyDeviceToHost, st));
    CP_HIP(hipMemcpyAsync(&nfail, dfail.get(), 4, hipMemcpyDeviceToHost, st));
    if (out->residual) {
        DevBuf d32(ctx, NP_ * 4);
        launch(k_f64_to_f32, dim3((unsigned)std::min<size_t>(cdiv((int64_t)NP_, BLOCK), 65536)), dim3(BLOCK), 0,
               st, (int64_t)NP_, (const double*)dR.as<double>(), d32.as<float>());
        CP_HIP(hipMemcpyAsync(out->residual, d32.get(), NP_ * 4, hipMemcpyDeviceToHost, st));
        CP_HIP(hipStreamSynchronize(st));
    }
    CP_HIP(hipStreamSynchronize(st));
    double* dcols[6] = {out->qfit, out->qfit_expected, out->qfit_z, out->crowding, out->spread_model,
                        out->spread_model_err};
    for (int i = 0; i < n; ++i) {
        const double* m = &prm[(size_t)i * npf];
        const double* e = &err[(size_t)i * npf];
        if (out->y) out->y[i] = m[0];
        if (out->x) out->x[i] = m[1];
        if (out->flux) out->flux[i] = m[rf];
        if (out->bkg) out->bkg[i] = m[rb];
        if (out->y_err) out->y_err[i] = e[0];
        if (out->x_err) out->x_err[i] = e[1];
        if (out->flux_err) out->flux_err[i] = e[rf];
        if (out->bkg_err) out->bkg_err[i] = e[rb];
        if (out->chisq) out->chisq[i] = chisq[i];
        if (out->converged) out->converged[i] = conv[i];
        if (out->valid) out->valid[i] = valid[i];
        if (out->n_iter) out->n_iter[i] = nit[i];
        for (int k = 0; k < 6; ++k)
            if (dcols[k]) dcols[k][i] = diag[(size_t)k * n + i];
    }
    out->n_passes = n_passes;
    out->n_failed = nfail;
}

}  // namespace
}  // namespace cp

using namespace cp;

extern "C" int cp_fit_all_stars(cp_ctx* ctx, const cp_psf_desc* psf, const float* image, const float* inv_var,
                                int32_t ny, int32_t nx, const double* theta0, int32_t n_stars, double fit_rad,
                                int32_t n_passes, uint32_t free_mask, const double* fixed_values,
                                const cp_lm_opts* opts, double spread_model_fwhm, int32_t precision,
                                cp_sim_result* out) {
    if (!ctx) return CP_ESTATE;
    return guarded(ctx, [&] {
        CP_HIP(hipSetDevice(ctx->device));
        fit_all_stars_impl(ctx, psf, image, inv_var, ny, nx, theta0, n_stars, fit_rad, n_passes, free_mask,
                           fixed_values, opts, spread_model_fwhm, precision, out);
    });
}
