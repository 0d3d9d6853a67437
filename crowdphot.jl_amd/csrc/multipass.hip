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

#include <hipcub/hipcub.hpp>

#include "cp_internal.h"
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
    __device__ __forceinline__ int rows() const { return by; }
    __device__ __forceinline__ void at(int m, int row, int col, double& py, double& px, double& v, double& w) const {
        py = (double)(ylo + row);
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

// Three waves per SIMD for the float32 fits of up to three free parameters
// (at most 24 bytes of spills, outside the per-pixel passes; the 3-pass fit of
// the 4096^2 crowded field 30.4 -> 29.9 ms), the compiler's choice otherwise.
template <typename T, int NP>
constexpr int mp_waves() {
    return sizeof(T) == 4 && NP <= 3 ? 3 : 1;
}

template <typename T, int KIND, int NP>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(mp_waves<T, NP>()))) void k_mp_level(MPArgs A, LMModel<T> Md, const int* __restrict__ list, int cnt) {
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
#pragma unroll
        for (int k = 0; k < 6; ++k)
            if (k < npf) A.prm[(size_t)s * npf + k] = R.full[k];
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
    const double flux = pick6(R.full, rf), bkg = pick6(R.full, rb);
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
    double sums[12] = {qfit, nclean, ndirty, dcrowd, Sp, Sq, Sp0, Sq0, SqG, ssum, s2sum, ngood};
    wave_sum_many<12>(sums);
    qfit = sums[0], nclean = sums[1], ndirty = sums[2], dcrowd = sums[3];
    Sp = sums[4], Sq = sums[5], Sp0 = sums[6], Sq0 = sums[7], SqG = sums[8];
    ssum = sums[9], s2sum = sums[10], ngood = sums[11];
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

// ---------------------------------------------------------------------------
// Per-pass schedule, built on the device
// ---------------------------------------------------------------------------
// Box of every valid star (floor(y - r):ceil(y + r) clamped to the image,
// psf_photometry_single.jl:353-357) and its flux-order key: an unsigned
// integer that orders like Julia's isless on the flux with NaN above
// everything, so a stable descending sort is sortperm(flux; rev = true).
// Stars whose box has fewer than 3 pixels become invalid (:358); invalid
// stars take key 0, below every valid key.  *bad = 1 + a star whose position
// is not finite (the reference's floor(Int, NaN) throws).
static __global__ void k_mp_boxes(int n, int npf, const double* __restrict__ prm, uint8_t* __restrict__ valid,
                                  double r, int ny, int nx, int4* __restrict__ box,
                                  unsigned long long* __restrict__ key, int* __restrict__ bad) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    key[i] = 0ull;
    if (!valid[i]) return;
    const double* m = prm + (size_t)i * npf;
    const double y = m[0], x = m[1];
    if (!(fabs(y) < 1e9 && fabs(x) < 1e9)) {
        *bad = i + 1;
        return;
    }
    const long ylo = max((long)floor(y - r), 1L), yhi = min((long)ceil(y + r), (long)ny);
    const long xlo = max((long)floor(x - r), 1L), xhi = min((long)ceil(x + r), (long)nx);
    const long by = yhi - ylo + 1, bx = xhi - xlo + 1;
    if (by <= 0 || bx <= 0 || by * bx < 3) {
        valid[i] = 0;
        return;
    }
    box[i] = make_int4((int)ylo, (int)yhi, (int)xlo, (int)xhi);
    const double f = m[npf - 2];
    const unsigned long long b = (unsigned long long)__double_as_longlong(f);
    key[i] = isnan(f) ? ~0ull : ((b >> 63) ? ~b : (b | (1ull << 63)));
}

// Cell of the star at flux rank k: the W x W cell of its box's lower corner
// (boxes are at most W pixels a side, so overlapping boxes sit in the same or
// adjacent cells); invalid ranks take the sentinel ncell.  Level 1 to start.
static __global__ void k_mp_cells(int n, const int* __restrict__ order, const unsigned long long* __restrict__ skey,
                                  const int4* __restrict__ box, int W, int ncy, int ncell, int* __restrict__ cell,
                                  int* __restrict__ rank_seq, int* __restrict__ lev) {
    const int k = blockIdx.x * BLOCK + threadIdx.x;
    if (k >= n) return;
    rank_seq[k] = k;
    if (skey[k] == 0ull) {
        cell[k] = ncell;
        lev[k] = 0;
        return;
    }
    const int4 b = box[order[k]];
    cell[k] = (b.x - 1) / W + ((b.z - 1) / W) * ncy;
    lev[k] = 1;
}

// Start / end of each cell's run in the cell-sorted rank list.
static __global__ void k_mp_cell_bounds(int n, const int* __restrict__ cs, int ncell, int* __restrict__ cstart,
                                        int* __restrict__ cend) {
    const int s = blockIdx.x * BLOCK + threadIdx.x;
    if (s >= n) return;
    const int c = cs[s];
    if (c >= ncell) return;
    if (s == 0 || cs[s - 1] != c) cstart[c] = s;
    if (s == n - 1 || cs[s + 1] != c) cend[c] = s + 1;
}

// One relaxation sweep of level(k) = 1 + max level of the earlier-ranked
// stars whose boxes overlap k's.  Levels only grow towards that longest-path
// value, so sweeps may read neighbours' levels from before or after their
// update in the same sweep; a sweep that changes nothing has reached it.
static __global__ void k_mp_relax(int n, const int* __restrict__ order, const int4* __restrict__ box,
                                  const int* __restrict__ cell, const int* __restrict__ cstart,
                                  const int* __restrict__ cend, const int* __restrict__ cranks, int ncy, int ncx,
                                  int* lev, int* __restrict__ changed) {
    const int k = blockIdx.x * BLOCK + threadIdx.x;
    if (k >= n) return;
    const int c = cell[k];
    if (c >= ncy * ncx) return;
    const int4 bj = box[order[k]];
    const int cy = c % ncy, cx = c / ncy;
    int best = 0;
    for (int dx = -1; dx <= 1; ++dx) {
        const int xx = cx + dx;
        if (xx < 0 || xx >= ncx) continue;
        for (int dy = -1; dy <= 1; ++dy) {
            const int yy = cy + dy;
            if (yy < 0 || yy >= ncy) continue;
            const int cc = yy + xx * ncy;
            const int e = cend[cc];
            for (int s = cstart[cc]; s < e; ++s) {
                const int r = cranks[s];
                if (r >= k) break;   // ranks ascend within a cell
                const int4 bi = box[order[r]];
                if (bi.y < bj.x || bj.y < bi.x || bi.w < bj.z || bj.w < bi.z) continue;
                best = max(best, __atomic_load_n(lev + r, __ATOMIC_RELAXED));
            }
        }
    }
    if (best + 1 > lev[k]) {
        __atomic_store_n(lev + k, best + 1, __ATOMIC_RELAXED);
        *changed = 1;
    }
}

// First position of each level in the level-sorted star list.
static __global__ void k_mp_level_bounds(int n, const int* __restrict__ ls, int* __restrict__ lstart) {
    const int s = blockIdx.x * BLOCK + threadIdx.x;
    if (s >= n) return;
    const int l = ls[s];
    if (l > 0 && (s == 0 || ls[s - 1] != l)) lstart[l] = s;
}

static __global__ void k_iota(int n, int* __restrict__ v) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) v[i] = i;
}

// Result columns from the parameter and error rows.
static __global__ void k_mp_results(int n, int npf, const double* __restrict__ prm, const double* __restrict__ err,
                                    double* __restrict__ cols) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const size_t N = (size_t)n;
    const double* m = prm + (size_t)i * npf;
    const double* e = err + (size_t)i * npf;
    cols[OUT_Y * N + i] = m[0];
    cols[OUT_X * N + i] = m[1];
    cols[OUT_FLUX * N + i] = m[npf - 2];
    cols[OUT_BKG * N + i] = m[npf - 1];
    cols[OUT_Y_ERR * N + i] = e[0];
    cols[OUT_X_ERR * N + i] = e[1];
    cols[OUT_FLUX_ERR * N + i] = e[npf - 2];
    cols[OUT_BKG_ERR * N + i] = e[npf - 1];
}

// Initial errors (NaN free, 0 fixed) and state.
static __global__ void k_mp_init(int n, int npf, uint32_t free_mask, double* __restrict__ err,
                                 uint8_t* __restrict__ valid, uint8_t* __restrict__ conv, int* __restrict__ n_iter,
                                 double* __restrict__ cols) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    for (int k = 0; k < npf; ++k) err[(size_t)i * npf + k] = ((free_mask >> k) & 1u) ? NAN : 0.0;
    valid[i] = 1;
    conv[i] = 0;
    n_iter[i] = 0;
    const size_t N = (size_t)n;
    cols[OUT_CHISQ * N + i] = 0.0;
    for (int k = OUT_QFIT; k < OUT_F64; ++k) cols[k * N + i] = NAN;
}

struct Sorter {
    cp_ctx* ctx;
    DevBuf tmp;
    void* get(size_t b) {
        if (tmp.bytes() < b) tmp.alloc(ctx, std::max<size_t>(b, 1 << 20));
        return tmp.get();
    }
};

void fit_all_stars_device(cp_ctx* ctx, const cp_psf_desc* psf, const float* d_image, const float* d_iv, int ny,
                          int nx, const double* d_theta0, int n, double fit_rad, int n_passes, uint32_t free_mask,
                          const double* fixed_values, const cp_lm_opts* o, double spread_model_fwhm, int precision,
                          cp_sim_result* out) {
    CP_ARG(psf && o && out && d_image, "NULL argument");
    CP_ARG(ny > 0 && nx > 0, "image must be non-empty");
    CP_ARG(n >= 0 && (d_theta0 || n == 0), "bad star catalogue");
    CP_ARG(n_passes > 0, "n_passes must be positive, got " + std::to_string(n_passes));
    CP_ARG(fit_rad > 0 && std::isfinite(fit_rad), "fit_rad must be positive");
    CP_ARG(precision == CP_PREC_MIXED || precision == CP_PREC_FP64, "unknown precision");
    LMConst C;
    int np = 0;
    const char* why = "";
    CP_ARG(lm_const(psf, free_mask, o, d_iv != nullptr, C, np, &why), why);
    const int npf = C.npf;
    const long W = 2 * (long)std::ceil(fit_rad) + 2;
    CP_ARG(W * W * MP_BYTES_PER_PIXEL <= MP_LDS_MAX, "fit_rad too large for the per-star box (at most " +
                                                         std::to_string((int)std::sqrt(MP_LDS_MAX / MP_BYTES_PER_PIXEL) / 2 - 1) + ")");
    out->n_failed = 0;
    if (n == 0) {   // the reference's empty result (:285-288)
        out->n_passes = 0;
        return;
    }
    const HostPsf HP = make_host_psf(psf);
    DevPsf DP;
    upload_psf(ctx, psf, HP, DP);
    const LMModel<double> Md64{model_const<double>(HP, DP, 0.0), DP.G};
    const LMModel<float> Md32{model_const<float>(HP, DP, 0.0), DP.G};
    const int maxb = (int)(W * W);
    const size_t lds = (size_t)maxb * MP_BYTES_PER_PIXEL;
    hipStream_t st = ctx->stream;
    const size_t NP_ = (size_t)ny * nx, N = (size_t)n;
    const unsigned gn = cdiv(n, BLOCK);

    DevBuf dR(ctx, NP_ * 8), dprm(ctx, N * npf * 8), derr(ctx, N * npf * 8), dbox(ctx, N * sizeof(int4)),
        dvalid(ctx, N), dconv(ctx, N), dnit(ctx, N * 4), dfail(ctx, 8), dcols(ctx, N * OUT_F64 * 8), dK;
    // schedule buffers
    const int Wi = (int)W, ncy = (ny + Wi - 1) / Wi, ncx = (nx + Wi - 1) / Wi, ncell = ncy * ncx;
    DevBuf dkey(ctx, N * 8), dskey(ctx, N * 8), diota(ctx, N * 4), dorder(ctx, N * 4), dcell(ctx, N * 4),
        dcell_s(ctx, N * 4), drank(ctx, N * 4), dcrank(ctx, N * 4), dlev(ctx, N * 4), dlev_s(ctx, N * 4),
        dlist(ctx, N * 4), dcstart(ctx, (size_t)ncell * 4), dcend(ctx, (size_t)ncell * 4),
        dlstart(ctx, (N + 2) * 4), dflag(ctx, 8);
    Sorter sorter{ctx, {}};
    launch(k_iota, dim3(gn), dim3(BLOCK), 0, st, n, diota.as<int>());
    launch(k_f32_to_f64, dim3((unsigned)std::min<size_t>(cdiv((int64_t)NP_, BLOCK), 65536)), dim3(BLOCK), 0, st,
           (int64_t)NP_, d_image, dR.as<double>());
    CP_HIP(hipMemcpyAsync(dprm.get(), d_theta0, N * npf * 8, hipMemcpyDeviceToDevice, st));
    launch(k_mp_init, dim3(gn), dim3(BLOCK), 0, st, n, npf, free_mask, derr.as<double>(), dvalid.as<uint8_t>(),
           dconv.as<uint8_t>(), dnit.as<int>(), dcols.as<double>());
    CP_HIP(hipMemsetAsync(dfail.get(), 0, 8, st));

    // spread_model reference disk, sized from star 1 (:320-323)
    const int kh = (int)std::ceil(fit_rad);
    double yx0[2];
    CP_HIP(hipMemcpyAsync(yx0, d_theta0, 16, hipMemcpyDeviceToHost, st));
    CP_HIP(hipStreamSynchronize(st));
    const double sfw = spread_model_fwhm > 0 ? spread_model_fwhm : spread_fwhm_host(psf, yx0[0], yx0[1]);
    std::vector<double> hK;
    if (std::isfinite(sfw) && sfw > 0) {
        hK.resize((size_t)(2 * kh + 1) * (2 * kh + 1));
        spread_kernel_host(sfw, kh, hK.data());
        dK.alloc(ctx, hK.size() * 8);
        CP_HIP(hipMemcpyAsync(dK.get(), hK.data(), hK.size() * 8, hipMemcpyHostToDevice, st));
    }

    MPArgs A{};
    A.C = C;
    A.free_mask = free_mask;
    for (int k = 0; k < 6; ++k) A.fixv[k] = (fixed_values && k < npf) ? fixed_values[k] : NAN;
    A.R = dR.as<double>();
    A.image = d_image;
    A.iv = d_iv;
    A.ny = ny;
    A.n = n;
    A.npf = npf;
    A.prm = dprm.as<double>();
    A.err = derr.as<double>();
    A.box = dbox.as<int4>();
    A.valid = dvalid.as<uint8_t>();
    A.conv = dconv.as<uint8_t>();
    A.chisq = dcols.as<double>() + OUT_CHISQ * N;
    A.n_iter = dnit.as<int>();
    A.n_failed = dfail.as<int>();
    A.K = hK.empty() ? nullptr : dK.as<double>();
    A.kh = kh;
    A.diag = dcols.as<double>() + OUT_QFIT * N;
    A.maxb = maxb;

    int cell_bits = 1;
    while ((1 << cell_bits) <= ncell) ++cell_bits;
    std::vector<int> lstart;
    for (int pass = 1; pass <= n_passes; ++pass) {
        // boxes, flux order (stable descending sort of the keys)
        CP_HIP(hipMemsetAsync(dflag.get(), 0, 8, st));
        launch(k_mp_boxes, dim3(gn), dim3(BLOCK), 0, st, n, npf, (const double*)A.prm, A.valid, fit_rad, ny, nx,
               dbox.as<int4>(), dkey.as<unsigned long long>(), dflag.as<int>() + 1);
        size_t b = 0;
        CP_HIP(hipcub::DeviceRadixSort::SortPairsDescending(nullptr, b, dkey.as<unsigned long long>(),
                                                            dskey.as<unsigned long long>(), diota.as<int>(),
                                                            dorder.as<int>(), n, 0, 64, st));
        CP_HIP(hipcub::DeviceRadixSort::SortPairsDescending(sorter.get(b), b, dkey.as<unsigned long long>(),
                                                            dskey.as<unsigned long long>(), diota.as<int>(),
                                                            dorder.as<int>(), n, 0, 64, st));
        // cells of the ordered stars, grouped by cell with ranks ascending
        launch(k_mp_cells, dim3(gn), dim3(BLOCK), 0, st, n, (const int*)dorder.as<int>(),
               (const unsigned long long*)dskey.as<unsigned long long>(), (const int4*)A.box, Wi, ncy, ncell,
               dcell.as<int>(), drank.as<int>(), dlev.as<int>());
        b = 0;
        CP_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, b, dcell.as<int>(), dcell_s.as<int>(), drank.as<int>(),
                                                  dcrank.as<int>(), n, 0, cell_bits, st));
        CP_HIP(hipcub::DeviceRadixSort::SortPairs(sorter.get(b), b, dcell.as<int>(), dcell_s.as<int>(),
                                                  drank.as<int>(), dcrank.as<int>(), n, 0, cell_bits, st));
        CP_HIP(hipMemsetAsync(dcstart.get(), 0, (size_t)ncell * 4, st));
        CP_HIP(hipMemsetAsync(dcend.get(), 0, (size_t)ncell * 4, st));
        launch(k_mp_cell_bounds, dim3(gn), dim3(BLOCK), 0, st, n, (const int*)dcell_s.as<int>(), ncell,
               dcstart.as<int>(), dcend.as<int>());
        int bad = 0;
        CP_HIP(hipMemcpyAsync(&bad, dflag.as<int>() + 1, 4, hipMemcpyDeviceToHost, st));
        CP_HIP(hipStreamSynchronize(st));
        CP_ARG(bad == 0, "star " + std::to_string(bad) + " has a non-finite or out-of-range position");
        // levels: relaxation sweeps until nothing changes
        for (int sweep = 0;; ++sweep) {
            CP_ARG(sweep <= n, "level relaxation did not settle");
            CP_HIP(hipMemsetAsync(dflag.get(), 0, 4, st));
            launch(k_mp_relax, dim3(gn), dim3(BLOCK), 0, st, n, (const int*)dorder.as<int>(), (const int4*)A.box,
                   (const int*)dcell.as<int>(), (const int*)dcstart.as<int>(), (const int*)dcend.as<int>(),
                   (const int*)dcrank.as<int>(), ncy, ncx, dlev.as<int>(), dflag.as<int>());
            int changed = 0;
            CP_HIP(hipMemcpyAsync(&changed, dflag.get(), 4, hipMemcpyDeviceToHost, st));
            CP_HIP(hipStreamSynchronize(st));
            if (!changed) break;
        }
        // stars grouped by level (level 0: invalid, first)
        int lbits = 1;
        while ((1 << lbits) <= n + 1) ++lbits;
        b = 0;
        CP_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, b, dlev.as<int>(), dlev_s.as<int>(), dorder.as<int>(),
                                                  dlist.as<int>(), n, 0, lbits, st));
        CP_HIP(hipcub::DeviceRadixSort::SortPairs(sorter.get(b), b, dlev.as<int>(), dlev_s.as<int>(),
                                                  dorder.as<int>(), dlist.as<int>(), n, 0, lbits, st));
        launch(k_mp_level_bounds, dim3(gn), dim3(BLOCK), 0, st, n, (const int*)dlev_s.as<int>(),
               dlstart.as<int>());
        int nlev = 0;
        CP_HIP(hipMemcpyAsync(&nlev, dlev_s.as<int>() + (n - 1), 4, hipMemcpyDeviceToHost, st));
        CP_HIP(hipStreamSynchronize(st));
        lstart.assign(nlev + 2, n);
        if (nlev > 0) {
            CP_HIP(hipMemcpyAsync(lstart.data() + 1, dlstart.as<int>() + 1, (size_t)nlev * 4, hipMemcpyDeviceToHost,
                                  st));
            CP_HIP(hipStreamSynchronize(st));
        }
        lstart[nlev + 1] = n;
        A.first = pass == 1;
        A.last = pass == n_passes;
        const bool f64 = precision == CP_PREC_FP64;
        for (int l = 1; l <= nlev; ++l) {
            const int* lst = dlist.as<int>() + lstart[l];
            const int cnt = lstart[l + 1] - lstart[l];
            if (cnt <= 0) continue;
            if (psf->kind == CP_PSF_CIRCULAR_GAUSSIAN) {
                if (f64) launch_level<double, KIND_GAUSS>(np, lds, st, A, Md64, lst, cnt);
                else launch_level<float, KIND_GAUSS>(np, lds, st, A, Md32, lst, cnt);
            } else if (psf->kind == CP_PSF_CIRCULAR_MOFFAT) {
                if (f64) launch_level<double, KIND_MOFFAT>(np, lds, st, A, Md64, lst, cnt);
                else launch_level<float, KIND_MOFFAT>(np, lds, st, A, Md32, lst, cnt);
            } else {
                if (f64) launch_level<double, KIND_EPSF>(np, lds, st, A, Md64, lst, cnt);
                else launch_level<float, KIND_EPSF>(np, lds, st, A, Md32, lst, cnt);
            }
        }
        launch(k_mp_validate, dim3(gn), dim3(BLOCK), 0, st, n, npf, (const double*)A.prm, (const uint8_t*)A.conv,
               A.valid);
    }

    // results: columns on the device, one staged read-back
    launch(k_mp_results, dim3(gn), dim3(BLOCK), 0, st, n, npf, (const double*)A.prm, (const double*)A.err,
           dcols.as<double>());
    double* host_cols[OUT_F64] = {out->y,        out->x,         out->flux,    out->bkg,     out->y_err,
                                  out->x_err,    out->flux_err,  out->bkg_err, out->chisq,   out->qfit,
                                  out->qfit_expected, out->qfit_z, out->crowding, out->spread_model,
                                  out->spread_model_err};
    std::vector<D2HSeg> segs;
    for (int k = 0; k < OUT_F64; ++k)
        if (host_cols[k]) segs.push_back({host_cols[k], dcols.as<double>() + k * N, N * 8});
    if (out->n_iter) segs.push_back({out->n_iter, dnit.get(), N * 4});
    if (out->converged) segs.push_back({out->converged, dconv.get(), N});
    if (out->valid) segs.push_back({out->valid, dvalid.get(), N});
    DevBuf d32;
    if (out->residual) {
        d32.alloc(ctx, NP_ * 4);
        launch(k_f64_to_f32, dim3((unsigned)std::min<size_t>(cdiv((int64_t)NP_, BLOCK), 65536)), dim3(BLOCK), 0,
               st, (int64_t)NP_, (const double*)dR.as<double>(), d32.as<float>());
        segs.push_back({out->residual, d32.get(), NP_ * 4});
    }
    int nfail = 0;
    segs.push_back({&nfail, dfail.get(), 4});
    staged_d2h(ctx, st, segs);
    out->n_passes = n_passes;
    out->n_failed = nfail;
}

}  // namespace
}  // namespace cp

using namespace cp;

extern "C" {

int cp_fit_all_stars_device(cp_ctx* ctx, const cp_psf_desc* psf, const float* d_image, const float* d_inv_var,
                            int32_t ny, int32_t nx, const double* d_theta0, int32_t n_stars, double fit_rad,
                            int32_t n_passes, uint32_t free_mask, const double* fixed_values,
                            const cp_lm_opts* opts, double spread_model_fwhm, int32_t precision,
                            cp_sim_result* out) {
    if (!ctx) return CP_ESTATE;
    return guarded(ctx, [&] {
        CP_HIP(hipSetDevice(ctx->device));
        fit_all_stars_device(ctx, psf, d_image, d_inv_var, ny, nx, d_theta0, n_stars, fit_rad, n_passes, free_mask,
                             fixed_values, opts, spread_model_fwhm, precision, out);
    });
}

int cp_fit_all_stars(cp_ctx* ctx, const cp_psf_desc* psf, const float* image, const float* inv_var, int32_t ny,
                     int32_t nx, const double* theta0, int32_t n_stars, double fit_rad, int32_t n_passes,
                     uint32_t free_mask, const double* fixed_values, const cp_lm_opts* opts,
                     double spread_model_fwhm, int32_t precision, cp_sim_result* out) {
    if (!ctx) return CP_ESTATE;
    return guarded(ctx, [&] {
        CP_ARG(image && ny > 0 && nx > 0 && n_stars >= 0 && (theta0 || n_stars == 0), "bad arguments");
        CP_HIP(hipSetDevice(ctx->device));
        hipStream_t st = ctx->stream;
        const size_t P = (size_t)ny * nx;
        const int npf = lm_nparams(psf ? psf->kind : 0);
        DevBuf dimg(ctx, P * 4), div, dth(ctx, (size_t)std::max(n_stars, 1) * npf * 8);
        CP_HIP(hipMemcpyAsync(dimg.get(), image, P * 4, hipMemcpyHostToDevice, st));
        if (inv_var) {
            div.alloc(ctx, P * 4);
            CP_HIP(hipMemcpyAsync(div.get(), inv_var, P * 4, hipMemcpyHostToDevice, st));
        }
        if (n_stars) CP_HIP(hipMemcpyAsync(dth.get(), theta0, (size_t)n_stars * npf * 8, hipMemcpyHostToDevice, st));
        fit_all_stars_device(ctx, psf, dimg.as<float>(), inv_var ? div.as<float>() : nullptr, ny, nx,
                             dth.as<double>(), n_stars, fit_rad, n_passes, free_mask, fixed_values, opts,
                             spread_model_fwhm, precision, out);
        CP_HIP(hipStreamSynchronize(st));
    });
}

}  // extern "C"
