// sim_kernels.h -- device kernels of the simultaneous Gauss-Newton fit
// (fit_all_stars_simultaneous, src/photometry/psf_photometry_simultaneous.jl).
//
// Design (MI355X, wave64):
//  * image tiles of TY x TX = 1024 px each own two star lists (stars whose
//    fit-box or model-box overlaps the tile), built once by a stable radix
//    sort so each list is in the reference's Morton star order.  Every
//    image-space product (model render, residual, cost, freezing
//    subtraction) is a per-pixel gather over the tile's list: no float
//    atomics, and each pixel sums its stars in the same order as the
//    reference's scatter-add.
//  * per-star products (fill + J^T r, steps, freeze tests) run one
//    wavefront per star with wave-shuffle reductions in float64; the inner
//    solve runs in star space on the normal matrix (normal_op.h).
//  * grid-wide sums: every block writes its partial with a plain store and a
//    one-block k_finish launch sums them in block order (deterministic run to
//    run).  No same-address atomics and no grid-wide fences: on gfx950 both
//    serialise at the L2 (about 12 ns per block).
//  * blocks stage the ePSF node grid in LDS (polyphase layout) once and walk
//    contiguous tile / star ranges.
#pragma once

#include "cp_internal.h"
#include "psf_device.h"

namespace cp {

constexpr int BLOCK = 256;
constexpr int NWAVE = BLOCK / 64;
constexpr int TILE_PX = 1024;           // pixels per tile
constexpr int PPT = TILE_PX / BLOCK;    // pixels per thread
constexpr int RBLOCK = 1024;            // render / fill blocks share one staged ePSF grid
constexpr int RWAVE = RBLOCK / 64;
constexpr int FBLOCK = 1024;            // fill blocks: two per CU with the ePSF grid in LDS
constexpr int FWAVE = FBLOCK / 64;
constexpr int MAX_EPL = 16;             // stamp entries per lane (S^2 <= 1024)
constexpr double DEPS = 2.220446049250313e-16;

enum : int { MODE_RESID = 0, MODE_COST = 1, MODE_FREEZE = 2, MODE_FINAL = 3 };

// Stamp values V: one row of VP = 4 values (the p <= 3 free columns, zero
// padded) per stamp pixel, so a row is a single 16-byte (float) load.
// b_scaled = J^T r (float64) keeps the same padded stride per star.
constexpr int VP = 4;
template <typename T> struct Vec4Of;
template <> struct Vec4Of<float> { using type = float4; };
template <> struct Vec4Of<double> { using type = double4; };
template <typename T> using vec4_t = typename Vec4Of<T>::type;

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2).  Renumber blocks so every XCD receives one contiguous range of logical
// blocks: neighbouring pixel / star ranges then share an L2.
__device__ __forceinline__ int xcd_remap(int bid, int nb) {
    constexpr int NX = 8;
    const int per = nb / NX, rem = nb % NX;
    const int x = bid % NX, i = bid / NX;
    return (x < rem ? x * (per + 1) : rem * (per + 1) + (x - rem) * per) + i;
}

// Device-resident scalars of one fit.
struct Scal {
    double cost, cost_cand, dof, bnorm2;
    unsigned long long gmax_bits;
    int n_live, n_nonfinite, n_frozen_now, pad0;
    unsigned long long n_union;
    // inner solver state
    double lam, beta1, alpha, beta_cur, beta_next, rhobar, phibar;
    double anorm2, dnorm2, xnorm2, c2, s2, z, res2, xenorm2, errv[5];
    double sig, tau;        // pending x/w update of the last finished iteration
    double atol, btol;
    int iter, done, itmax, solver;
    // stars live at the last compaction (the lists the fill, H and the star
    // vectors walk) and the row entries of their rows (the Golub-Kahan step's)
    int n_jtu_stars, n_gk_entries;
    // CG on the damped normal equations: gamma = |s|^2, pp = |p|^2, the step a
    // and the direction update beta of the current iteration, |s_0|
    double gamma, cg_alpha, cg_beta, pnorm2, qnorm2, s0norm;
    // LSMR (Fong & Saunders 2011): the QR rotations, the ||r|| estimate, the
    // cond(A) bounds, ||x_k||^2, and del, the hbar factor of the pending
    // update (0 for LSQR, whose pending update is LSMR's with del = 0)
    double alphabar, rho, rhobar_m, cbar, sbar, zeta, zetabar;
    double betadd, betad, rhodold, tautildeold, thetatilde, dsum, maxrbar, minrbar, del;
    unsigned int counters[8];
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Cross-lane moves of float64 values without LDS (gfx950): the permlane
// swaps exchange the upper 32 lanes of one register with the lower 32 of
// another (rows 1, 3 with rows 0, 2 for the 16-lane form); DPP reads lane
// l ^ 8 (row rotate by 8), l +- 4 (row shifts) and lanes of the same quad.
__device__ __forceinline__ unsigned f64_lo(double x) {
    return (unsigned)(unsigned long long)__double_as_longlong(x);
}
__device__ __forceinline__ unsigned f64_hi(double x) {
    return (unsigned)((unsigned long long)__double_as_longlong(x) >> 32);
}
__device__ __forceinline__ double f64_of(unsigned lo, unsigned hi) {
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// (a, b) -> a = [a lanes 0-31, b lanes 0-31], b = [a lanes 32-63, b lanes 32-63]
// (D = 32); rows of 16 for D = 16: a = [a r0, b r0, a r2, b r2], b = [a r1, b r1,
// a r3, b r3].
template <int D>
__device__ __forceinline__ void lane_swap_f64(double& a, double& b) {
    static_assert(D == 32 || D == 16, "lane_swap_f64: D = 32 or 16");
    if constexpr (D == 32) {
        const auto l = __builtin_amdgcn_permlane32_swap(f64_lo(a), f64_lo(b), false, false);
        const auto h = __builtin_amdgcn_permlane32_swap(f64_hi(a), f64_hi(b), false, false);
        a = f64_of(l[0], h[0]);
        b = f64_of(l[1], h[1]);
    } else {
        const auto l = __builtin_amdgcn_permlane16_swap(f64_lo(a), f64_lo(b), false, false);
        const auto h = __builtin_amdgcn_permlane16_swap(f64_hi(a), f64_hi(b), false, false);
        a = f64_of(l[0], h[0]);
        b = f64_of(l[1], h[1]);
    }
}

template <int CTRL>
__device__ __forceinline__ double dpp_f64(double x) {
    return f64_of((unsigned)__builtin_amdgcn_update_dpp(0, (int)f64_lo(x), CTRL, 0xf, 0xf, true),
                  (unsigned)__builtin_amdgcn_update_dpp(0, (int)f64_hi(x), CTRL, 0xf, 0xf, true));
}

// The value lane l ^ D holds (D = 8, 4, 2, 1), every lane reading
// unconditionally (a select between two DPP reads is compiled as a branch,
// and a lane whose source lane is masked off reads 0).  l ^ 4 is the mirror
// within 8 lanes (l ^ 7) followed by the mirror within the quad (^ 3).
template <int D>
__device__ __forceinline__ double lane_xor_f64(double x) {
    if constexpr (D == 8) return dpp_f64<0x128>(x);                 // row_ror:8
    else if constexpr (D == 4) return dpp_f64<0x1B>(dpp_f64<0x141>(x));  // row_half_mirror, quad_perm [3,2,1,0]
    else if constexpr (D == 2) return dpp_f64<0x4E>(x);             // quad_perm [2,3,0,1]
    else return dpp_f64<0xB1>(x);                                   // quad_perm [1,0,3,2]
}

// One reduce-scatter step across lane bit D over the first 2 HALF values:
// every index is a compile-time constant, so the values stay in registers
// (a select between array elements at a loop index becomes a dynamically
// indexed array otherwise).  Across bits 5 and 4 the swap leaves the kept
// half in one register and the partner's copy of it in the other.
template <int HALF, int D, int KP, size_t... I>
__device__ __forceinline__ void wave_rs_step(double (&w)[KP], int lane, std::index_sequence<I...>) {
    if constexpr (D >= 16) {
        ((lane_swap_f64<D>(w[I], w[I + HALF]), w[I] = w[I] + w[I + HALF]), ...);
    } else {
        const bool hi = (lane & D) != 0;
        ((w[I] = (hi ? w[I + HALF] : w[I]) + lane_xor_f64<D>(hi ? w[I] : w[I + HALF])), ...);
    }
}

template <int HALF, int D, int KP>
__device__ __forceinline__ void wave_rs(double (&w)[KP], int lane) {
    if constexpr (HALF >= 1) {
        wave_rs_step<HALF, D>(w, lane, std::make_index_sequence<HALF>{});
        wave_rs<HALF / 2, D / 2>(w, lane);
    }
}

// t + (the value lane l ^ D holds), for D = 32 .. 1
template <int D>
__device__ __forceinline__ double lane_pair_sum(double t, int lane) {
    if constexpr (D >= 16) {
        double u = t;
        lane_swap_f64<D>(t, u);
        return t + u;
    } else {
        return t + lane_xor_f64<D>(t);
    }
}

template <int D>
__device__ __forceinline__ double wave_bfly(double t, int lane) {
    if constexpr (D >= 1) return wave_bfly<D / 2>(lane_pair_sum<D>(t, lane), lane);
    else return t;
}

// All-gather step across lane bit D: each lane holds CNT values and ends with
// its own and its partner's, the half of the lane with bit D clear first.
template <int CNT, int D, int KP, size_t... I>
__device__ __forceinline__ void wave_ag_step(double (&w)[KP], int lane, std::index_sequence<I...>) {
    if constexpr (D >= 16) {
        ((w[I + CNT] = w[I], lane_swap_f64<D>(w[I], w[I + CNT])), ...);
    } else {
        const bool hi = (lane & D) != 0;
        double r[CNT];
        ((r[I] = lane_xor_f64<D>(w[I])), ...);
        ((w[I + CNT] = hi ? w[I] : r[I], w[I] = hi ? r[I] : w[I]), ...);
    }
}

template <int CNT, int D, int KP>
__device__ __forceinline__ void wave_ag(double (&w)[KP], int lane) {
    if constexpr (CNT < KP) {
        wave_ag_step<CNT, D>(w, lane, std::make_index_sequence<CNT>{});
        wave_ag<CNT * 2, D * 2>(w, lane);
    }
}

// Wave sums of K float64 values at once, returned to every lane.  A
// reduce-scatter over lane bits 5 .. 6-log2(KP) (KP = K rounded up to a power
// of two): at each step a lane keeps half of its values, sends the other half
// to the partner lane across that bit and adds the partner's copy of the half
// it kept; the one value left is summed over the remaining lane bits, lanes
// (64/KP)i .. (64/KP)(i+1)-1 then hold total i, and an all-gather over the
// reduce-scatter's bits in reverse hands every lane all K totals.  Each
// partial sum adds a lane's value to its partner's across the same bits in the
// same order as wave_sum (32, 16, ..., 1), so every total is wave_sum's bit
// for bit.  The moves are permlane swaps and DPP, no LDS.  The whole wave
// must be converged.
template <int K>
__device__ __forceinline__ void wave_sum_many(double (&v)[K]) {
    static_assert(K >= 1 && K <= 32, "wave_sum_many: 1..32 values");
    constexpr int KP = K <= 1 ? 1 : K <= 2 ? 2 : K <= 4 ? 4 : K <= 8 ? 8 : K <= 16 ? 16 : 32;
    const int lane = threadIdx.x & 63;
    double w[KP];
#pragma unroll
    for (int i = 0; i < KP; ++i) w[i] = i < K ? v[i] : 0.0;
    wave_rs<KP / 2, 32>(w, lane);
    w[0] = wave_bfly<32 / KP>(w[0], lane);
    wave_ag<1, 64 / KP>(w, lane);
#pragma unroll
    for (int i = 0; i < K; ++i) v[i] = w[i];
}

__device__ __forceinline__ int wave_sum_int(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// Block sum in fixed order over NW waves; result valid in thread 0.
template <int NW = NWAVE>
__device__ __forceinline__ double block_sum(double v, double* sh) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double r = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < NW; ++i) r += sh[i];
    return r;
}

// Block-sum v and store it as this block's partial (plain store; the k_finish
// launch that follows reads it after the kernel boundary).
template <int NW = NWAVE>
__device__ __forceinline__ void block_store(double v, double* partials, double* sh) {
    const double b = block_sum<NW>(v, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = b;
}

template <int NW = NWAVE>
__device__ __forceinline__ double block_max(double v, double* sh) {
    v = wave_max(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double r = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < NW; ++i) r = fmax(r, sh[i]);
    return r;
}

// Symmetric Givens rotation: [c s; s -c] [a; b] = [rho; 0].
__device__ __forceinline__ void sym_givens(double a, double b, double& c, double& s, double& rho) {
    auto sgn = [](double v) { return v > 0 ? 1.0 : (v < 0 ? -1.0 : 0.0); };
    if (b == 0) {
        c = (a == 0) ? 1.0 : sgn(a);
        s = 0;
        rho = fabs(a);
    } else if (a == 0) {
        c = 0;
        s = sgn(b);
        rho = fabs(b);
    } else if (fabs(b) > fabs(a)) {
        double t = a / b;
        double tau = sgn(b) * sqrt(1 + t * t);
        s = 1 / tau;
        c = s * t;
        rho = b * tau;
    } else {
        double t = b / a;
        double tau = sgn(a) * sqrt(1 + t * t);
        c = 1 / tau;
        s = c * t;
        rho = a * tau;
    }
}

// ---------------------------------------------------------------------------
// setup kernels
// ---------------------------------------------------------------------------

// Data and weights (psf_photometry_simultaneous.jl:1056-1084).
template <typename T>
static __global__ void k_prep_pixels(const float* __restrict__ img, const float* __restrict__ iv, int64_t P,
                              T* __restrict__ data, T* __restrict__ dw, T* __restrict__ w, Scal* S) {
    int bad = 0;
    for (int64_t q = blockIdx.x * (int64_t)BLOCK + threadIdx.x; q < P; q += (int64_t)gridDim.x * BLOCK) {
        float d = img[q];
        float wv = 1.0f;
        if (iv) {
            float t = iv[q];
            wv = (isfinite(t) && t > 0.0f) ? t : 0.0f;
        }
        if (!isfinite(d)) {
            bad += (iv == nullptr);
            d = 0.0f;
            wv = 0.0f;
        }
        data[q] = (T)d;
        dw[q] = (T)d;
        w[q] = (T)wv;
    }
    if (bad) atomicAdd(&S->n_nonfinite, bad);
}

__device__ __forceinline__ uint64_t spread32(uint64_t v) {
    v &= 0x00000000ffffffffULL;
    v = (v | (v << 16)) & 0x0000ffff0000ffffULL;
    v = (v | (v << 8)) & 0x00ff00ff00ff00ffULL;
    v = (v | (v << 4)) & 0x0f0f0f0f0f0f0f0fULL;
    v = (v | (v << 2)) & 0x3333333333333333ULL;
    v = (v | (v << 1)) & 0x5555555555555555ULL;
    return v;
}

// Anchors (round half to even), activity and Morton key per catalogue star
// (_build_stamps!, :245-306; _morton2d :163-174).
template <typename T>
static __global__ void k_anchor(const double* __restrict__ y0, const double* __restrict__ x0,
                         const double* __restrict__ f0, int n, const T* __restrict__ w, int ny, int nx,
                         int R, int* __restrict__ ay, int* __restrict__ ax, uint8_t* __restrict__ ok) {
    int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    double y = y0[i], x = x0[i];
    bool fy = isfinite(y) && fabs(rint(y)) <= 1073741824.0;
    bool fx = isfinite(x) && fabs(rint(x)) <= 1073741824.0;
    int a = fy ? (int)rint(y) : INT_MIN;
    int b = fx ? (int)rint(x) : INT_MIN;
    ay[i] = a;
    ax[i] = b;
    bool any = false;
    if (fy && fx) {
        for (int dx = -R; dx <= R && !any; ++dx)
            for (int dy = -R; dy <= R; ++dy) {
                int gy = a + dy, gx = b + dx;
                if (gy >= 1 && gy <= ny && gx >= 1 && gx <= nx &&
                    w[(size_t)(gy - 1) + (size_t)(gx - 1) * ny] > T(0)) {
                    any = true;
                    break;
                }
            }
    }
    ok[i] = any && isfinite(f0[i]) && a >= 0 && b >= 0;
}

static __global__ void k_morton_keys(const int* __restrict__ sel, int na, const int* __restrict__ ay,
                              const int* __restrict__ ax, unsigned long long* __restrict__ keys) {
    int j = blockIdx.x * BLOCK + threadIdx.x;
    if (j >= na) return;
    int i = sel[j];
    keys[j] = spread32((uint64_t)ay[i]) | (spread32((uint64_t)ax[i]) << 1);
}

// Per active star (Morton order): anchors, theta, model half-width, state.
static __global__ void k_star_setup(const int* __restrict__ act, int na, const int* __restrict__ ay_cat,
                             const int* __restrict__ ax_cat, const double* __restrict__ y0,
                             const double* __restrict__ x0, const double* __restrict__ f0, int p, int gc0,
                             int gc1, int gc2, const int* __restrict__ mR_src, int mR_scalar,
                             int* __restrict__ ay, int* __restrict__ ax, int* __restrict__ mR,
                             double* __restrict__ theta, uint8_t* __restrict__ live,
                             int* __restrict__ frozen_at, uint8_t* __restrict__ conv) {
    int a = blockIdx.x * BLOCK + threadIdx.x;
    if (a >= na) return;
    int i = act[a];
    ay[a] = ay_cat[i];
    ax[a] = ax_cat[i];
    mR[a] = mR_src ? mR_src[a] : mR_scalar;
    const int gc[3] = {gc0, gc1, gc2};
    for (int k = 0; k < p; ++k) {
        double v = gc[k] == 1 ? y0[i] : (gc[k] == 2 ? x0[i] : f0[i]);
        theta[(size_t)a * p + k] = v;
    }
    live[a] = 1;
    frozen_at[a] = 0;
    conv[a] = 0;
}

// union_pix as a byte mask (_touched_pixels, :329-335).
template <typename T>
static __global__ void k_union(int na, const int* __restrict__ ay, const int* __restrict__ ax, int R, int S,
                        const T* __restrict__ w, int ny, int nx, uint8_t* __restrict__ umask) {
    const int S2 = S * S;
    int64_t e = blockIdx.x * (int64_t)BLOCK + threadIdx.x;
    if (e >= (int64_t)na * S2) return;
    int a = (int)(e / S2), m = (int)(e % S2);
    int gy = ay[a] + (m % S) - R, gx = ax[a] + (m / S) - R;
    if (gy < 1 || gy > ny || gx < 1 || gx > nx) return;
    size_t q = (size_t)(gy - 1) + (size_t)(gx - 1) * ny;
    if (w[q] > T(0)) umask[q] = 1;
}

static __global__ void k_count_mask(const uint8_t* __restrict__ m, int64_t P, double* __restrict__ partials) {
    __shared__ double sred[NWAVE];
    int c = 0;
    const int64_t P4 = P / 4;
    const uint32_t* m4 = reinterpret_cast<const uint32_t*>(m);
    for (int64_t q = blockIdx.x * (int64_t)BLOCK + threadIdx.x; q < P4; q += (int64_t)gridDim.x * BLOCK) {
        const uint32_t v = m4[q];
        c += (v & 0xff) + ((v >> 8) & 0xff) + ((v >> 16) & 0xff) + (v >> 24);
    }
    if (blockIdx.x == 0 && threadIdx.x < (int)(P - P4 * 4)) c += m[P4 * 4 + threadIdx.x];
    block_store((double)c, partials, sred);
}

// Stamp pixel indices of active stars (1-based, 0 = masked), for export.
template <typename T>
static __global__ void k_pixels(int na, const int* __restrict__ ay, const int* __restrict__ ax, int R, int S,
                         const T* __restrict__ w, int ny, int nx, int* __restrict__ pix) {
    const int S2 = S * S;
    int64_t e = blockIdx.x * (int64_t)BLOCK + threadIdx.x;
    if (e >= (int64_t)na * S2) return;
    int a = (int)(e / S2), m = (int)(e % S2);
    int gy = ay[a] + (m % S) - R, gx = ax[a] + (m / S) - R;
    int v = 0;
    if (gy >= 1 && gy <= ny && gx >= 1 && gx <= nx) {
        size_t q = (size_t)(gy - 1) + (size_t)(gx - 1) * ny;
        if (w[q] > T(0)) v = (int)(q + 1);
    }
    pix[e] = v;
}

// ---------------------------------------------------------------------------
// tile binning
// ---------------------------------------------------------------------------
struct TileGeo {
    int TY, TX, nty, ntx, ny, nx;
};

__device__ __forceinline__ bool tile_range(const TileGeo& g, int ay, int ax, int h, int& ty0, int& ty1,
                                           int& tx0, int& tx1) {
    int y0 = max(ay - h, 1), y1 = min(ay + h, g.ny);
    int x0 = max(ax - h, 1), x1 = min(ax + h, g.nx);
    if (y0 > y1 || x0 > x1) return false;
    ty0 = (y0 - 1) / g.TY;
    ty1 = (y1 - 1) / g.TY;
    tx0 = (x0 - 1) / g.TX;
    tx1 = (x1 - 1) / g.TX;
    return true;
}

static __global__ void k_tile_count(int na, const int* __restrict__ ay, const int* __restrict__ ax,
                             const int* __restrict__ hw, int h_scalar, TileGeo g, int* __restrict__ cnt) {
    int a = blockIdx.x * BLOCK + threadIdx.x;
    if (a >= na) return;
    int h = hw ? hw[a] : h_scalar, ty0, ty1, tx0, tx1;
    cnt[a] = tile_range(g, ay[a], ax[a], h, ty0, ty1, tx0, tx1) ? (ty1 - ty0 + 1) * (tx1 - tx0 + 1) : 0;
}

static __global__ void k_tile_emit(int na, const int* __restrict__ ay, const int* __restrict__ ax,
                            const int* __restrict__ hw, int h_scalar, TileGeo g, const int* __restrict__ off,
                            unsigned int* __restrict__ keys, int* __restrict__ vals) {
    int a = blockIdx.x * BLOCK + threadIdx.x;
    if (a >= na) return;
    int h = hw ? hw[a] : h_scalar, ty0, ty1, tx0, tx1;
    if (!tile_range(g, ay[a], ax[a], h, ty0, ty1, tx0, tx1)) return;
    int o = off[a];
    for (int tx = tx0; tx <= tx1; ++tx)
        for (int ty = ty0; ty <= ty1; ++ty) {
            keys[o] = (unsigned int)(ty + tx * g.nty);
            vals[o] = a;
            ++o;
        }
}

static __global__ void k_tile_ranges(const unsigned int* __restrict__ keys, int npairs, int* __restrict__ tstart,
                              int* __restrict__ tend) {
    int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= npairs) return;
    unsigned int k = keys[i];
    if (i == 0 || keys[i - 1] != k) tstart[k] = i;
    if (i == npairs - 1 || keys[i + 1] != k) tend[k] = i + 1;
}

// ---------------------------------------------------------------------------
// star records
// ---------------------------------------------------------------------------
struct FreeMap {
    int p, gc0, gc1, gc2;   // gradient column (1 = y, 2 = x, 3 = flux) of free slot k
    int y_fixed, x_fixed, f_fixed;
    double fixed_y, fixed_x, fixed_f;
};

__host__ __device__ __forceinline__ int gc_of(const FreeMap& fm, int k) {
    return k == 0 ? fm.gc0 : (k == 1 ? fm.gc1 : fm.gc2);
}

__device__ __forceinline__ void star_yxf(const FreeMap& fm, const double* th, double& Y, double& X,
                                         double& F) {
    Y = fm.fixed_y;
    X = fm.fixed_x;
    F = fm.fixed_f;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (k < fm.p) {
            const int c = gc_of(fm, k);
            const double v = th[k];
            if (c == 1) Y = v;
            if (c == 2) X = v;
            if (c == 3) F = v;
        }
    }
}

struct NodeGrid {
    int ngy, ngx;
    const double* ygrid;
    const double* xgrid;
    const int* index_grid;   // ngy x ngx column-major, 0-based node ids
};

// _grid_corners_dw (gridded_psf.jl:176-208).
template <typename T>
__device__ void grid_corners(const NodeGrid& G, int n_nodes, double Y, double X, CornerRec<T>& c) {
    if (n_nodes == 1) {
        c.idx[0] = c.idx[1] = c.idx[2] = c.idx[3] = 0;
        c.w[0] = T(1);
        c.w[1] = c.w[2] = c.w[3] = T(0);
        for (int k = 0; k < 4; ++k) c.dwdy[k] = c.dwdx[k] = T(0);
        return;
    }
    auto last_le = [](const double* g, int n, double v) {
        int lo = 0, hi = n;  // count of elements <= v
        while (lo < hi) {
            int mid = (lo + hi) >> 1;
            if (g[mid] <= v) lo = mid + 1; else hi = mid;
        }
        return lo;
    };
    int iy = min(max(last_le(G.ygrid, G.ngy, Y), 1), G.ngy - 1);
    int ix = min(max(last_le(G.xgrid, G.ngx, X), 1), G.ngx - 1);
    double y0 = G.ygrid[iy - 1], y1 = G.ygrid[iy], x0 = G.xgrid[ix - 1], x1 = G.xgrid[ix];
    double Yc = fmin(fmax(Y, y0), y1), Xc = fmin(fmax(X, x0), x1);
    double cy = (Y > y0 && Y < y1) ? 1.0 : 0.0, cx = (X > x0 && X < x1) ? 1.0 : 0.0;
    double inv = 1.0 / ((x1 - x0) * (y1 - y0));
    double dy1 = y1 - Yc, dy0 = Yc - y0, dx1 = x1 - Xc, dx0 = Xc - x0;
    auto ig = [&](int a, int b) { return G.index_grid[(a - 1) + (b - 1) * G.ngy]; };
    c.idx[0] = ig(iy, ix);
    c.idx[1] = ig(iy, ix + 1);
    c.idx[2] = ig(iy + 1, ix);
    c.idx[3] = ig(iy + 1, ix + 1);
    double w[4] = {dx1 * dy1 * inv, dx0 * dy1 * inv, dx1 * dy0 * inv, dx0 * dy0 * inv};
    double wy[4] = {-dx1 * cy * inv, -dx0 * cy * inv, dx1 * cy * inv, dx0 * cy * inv};
    double wx[4] = {-dy1 * cx * inv, dy1 * cx * inv, -dy0 * cx * inv, dy0 * cx * inv};
    for (int k = 0; k < 4; ++k) {
        c.w[k] = (T)w[k];
        c.dwdy[k] = (T)wy[k];
        c.dwdx[k] = (T)wx[k];
    }
}

// StarRec (and the single-node ePSF constants) from theta, for the stars of
// lst (Sc->n_jtu_stars of them) or, with lst null, for all.  frozen_iter >= 1
// marks stars frozen at that iteration.
template <typename T, int KIND>
static __global__ void k_star_rec(int na, const int* __restrict__ lst, const Scal* Sc,
                                  const double* __restrict__ theta, FreeMap fm, ModelConst<T> M,
                                  double inv_norm_g64, double moffat_amp64, const int* __restrict__ ay,
                                  const int* __restrict__ ax, const int* __restrict__ mR,
                                  const uint8_t* __restrict__ live, const int* __restrict__ frozen_at,
                                  int frozen_iter, int all_live, StarRec<T>* __restrict__ recs, NodeGrid G,
                                  CornerRec<T>* __restrict__ crs, EpsfRec<T>* __restrict__ ers) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= (lst ? Sc->n_jtu_stars : na)) return;
    // the freeze render's records: nothing to do when no star froze
    if (frozen_iter > 0 && Sc->n_frozen_now == 0) return;
    const int a = lst ? lst[i] : i;
    double Y, X, F;
    star_yxf(fm, theta + (size_t)a * fm.p, Y, X, F);
    StarRec<T> r;
    r.ay = ay[a];
    r.ax = ax[a];
    const double fy = Y - r.ay, fx = X - r.ax;
    r.fy = (T)fy;
    r.fx = (T)fx;
    r.flux = (T)F;
    if (KIND == KIND_GAUSS) r.amp = (T)(F * inv_norm_g64);
    else if (KIND == KIND_MOFFAT) r.amp = (T)(F * moffat_amp64);
    else r.amp = (T)F;
    r.mR = mR[a];
    int fl = (all_live || live[a]) ? 1 : 0;
    if (frozen_iter > 0 && frozen_at[a] == frozen_iter && !live[a]) fl |= 2;
    r.flags = fl;
    recs[a] = r;
    if (KIND == KIND_EPSF) {
        if (M.n_nodes > 1) {
            grid_corners<T>(G, M.n_nodes, Y, X, crs[a]);
        } else if (ers) {
            EpsfRec<T> e;   // built in registers, stored whole
            epsf_rec<T>(M, fy, fx, e);
            ers[a] = e;
        }
    }
}

// Tiles touched by the model box of a star of lst that froze at iteration
// iter get mark iter (plain stores of the same value; marks of earlier
// iterations need no clearing): the freeze render visits only those.
static __global__ void k_mark_frozen(const int* __restrict__ lst, const Scal* Sc, const uint8_t* __restrict__ live,
                                     const int* __restrict__ frozen_at, int iter, const int* __restrict__ ay,
                                     const int* __restrict__ ax, const int* __restrict__ mR, TileGeo g,
                                     int* __restrict__ tmark) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= Sc->n_jtu_stars) return;
    const int a = lst[i];
    if (live[a] || frozen_at[a] != iter) return;
    int ty0, ty1, tx0, tx1;
    if (!tile_range(g, ay[a], ax[a], mR[a], ty0, ty1, tx0, tx1)) return;
    for (int tx = tx0; tx <= tx1; ++tx)
        for (int ty = ty0; ty <= ty1; ++ty) tmark[tx * g.nty + ty] = iter;
}

// The device scalars copied to pinned host memory by the threads of one block
// (vector stores; the host reads them after a stream synchronize).
__device__ __forceinline__ void mirror_scal(const Scal* __restrict__ Sc, Scal* __restrict__ hS) {
    static_assert(sizeof(Scal) % 8 == 0, "Scal copied as 8-byte words");
    const unsigned long long* s = reinterpret_cast<const unsigned long long*>(Sc);
    unsigned long long* d = reinterpret_cast<unsigned long long*>(hS);
    for (int i = threadIdx.x; i < (int)(sizeof(Scal) / 8); i += blockDim.x) d[i] = s[i];
}
static __global__ void k_mirror(const Scal* __restrict__ Sc, Scal* __restrict__ hS) { mirror_scal(Sc, hS); }

// ---------------------------------------------------------------------------
// tile-gather render: residual / candidate cost / freeze subtraction / final
// (_render_model! :545-673, _residual_cost! :675-684, _cost! :686-693,
//  _freeze_star! :739-754, final render :1344-1356)
// ---------------------------------------------------------------------------
// Render tiles are RT x RT pixels, RT = 64 (the default) or 32, chosen per
// context (cp_tuning.render_tile); the 16 waves of a block own its 4 x 4
// sub-blocks of side RT / 4.
constexpr int RT = 64;      // default render tile side (rows = columns)
static_assert(RWAVE == 16, "a render tile is 4 x 4 sub-blocks, one per wave");

// tile-gather render.  Each 1024-thread block walks a contiguous range of
// RT x RT tiles; wave w owns the sub-block (w % 4, w / 4) of side SB = RT / 4
// and each lane the pixels (row lane % SB, columns (64 / SB) k + lane / SB),
// k < SB^2 / 64.  Every wave walks the tile's star list in chunks of 64, one
// entry per lane: the lane reads its star's record, tests whether this mode
// renders the star (live / newly frozen / all) and whether its model box
// touches the wave's sub-block, and prepares the star's clipped box and
// stencil constants; the wave then takes the touching stars in list order
// (ballot) and reads each one's constants from its lane (readlane).  No block
// barrier: the 16 waves of a tile run independently.  Per-pixel sums run over
// stars in list (Morton) order: deterministic and the reference's summation
// order.
// Render cost of a tile: its star-list length plus a fixed per-tile term.
constexpr int TILE_COST0 = 16;
static __global__ void k_tile_cost(const int* __restrict__ tstart, const int* __restrict__ tend, int ntiles,
                                   long long* __restrict__ cost) {
    const int t = blockIdx.x * BLOCK + threadIdx.x;
    if (t < ntiles) cost[t] = (long long)(tend[t] - tstart[t]) + TILE_COST0;
}

// Tiles [t0, t1) of logical block b of nb: equal shares of the inclusive cost
// scan (tiles whose scan value falls in (total b / nb, total (b + 1) / nb]).
__device__ __forceinline__ void tile_share(const long long* __restrict__ scan, int ntiles, int b, int nb, int& t0,
                                           int& t1) {
    const long long total = scan[ntiles - 1];
    auto upper = [&](long long v) {   // number of tiles with scan <= v
        int lo = 0, hi = ntiles;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (scan[mid] <= v) lo = mid + 1;
            else hi = mid;
        }
        return lo;
    };
    t0 = upper(total * b / nb);
    t1 = upper(total * (b + 1) / nb);
}

// Every render block's share of the tiles, computed once per cost scan (the
// scan changes only when the tile lists are rebuilt), so a render block
// starts with one load instead of two binary searches.
static __global__ void k_tile_shares(const long long* __restrict__ scan, int ntiles, int nb,
                                     int2* __restrict__ shares) {
    const int b = blockIdx.x * BLOCK + threadIdx.x;
    if (b >= nb) return;
    int t0, t1;
    tile_share(scan, ntiles, b, nb, t0, t1);
    shares[b] = make_int2(t0, t1);
}

// ceil(65536 / h) for h = 1..16: (p * RENDER_INV_H[h]) >> 16 == p / h for 0 <= p < 256
__constant__ const int RENDER_INV_H[17] = {0,    65536, 32768, 21846, 16384, 13108, 10923, 9363, 8192,
                                           7282, 6554,  5958,  5462,  5042,  4682,  4370,  4096};

// Dynamic LDS of the render: the staged ePSF grid (ngrid floats), then per
// wave an SB x SB accumulator of T and an SB x SB byte touch map.
template <typename T>
__host__ __device__ __forceinline__ size_t render_acc_offset(int ngrid) {
    return ((size_t)ngrid * 4 + 15) & ~(size_t)15;
}
template <typename T>
__host__ __device__ __forceinline__ size_t render_lds_bytes(int ngrid, int rt) {
    const size_t sb = (size_t)rt / 4;
    return render_acc_offset<T>(ngrid) + (size_t)RWAVE * sb * sb * (sizeof(T) + 1);
}

// Lane j's value of v (uniform: readlane), for float and double.
__device__ __forceinline__ float lane_value(float v, int j) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
}
__device__ __forceinline__ double lane_value(double v, int j) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(unsigned)(b & 0xffffffffll), j);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), j);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// Planes of the renders' fixed-layout ePSF path: os_y = os_x = 4 grids whose
// polyphase planes fit RQF x RQF cells are staged with planes of exactly that
// side, so every tap offset is a constant.
constexpr int RQF = 28;

// The pixels of one star's clipped box (n of them, h rows, local rows /
// columns from r_lo / c_lo) for a star whose whole box lies in its in-range
// window, phase PH and planes of QF x QF cells: the 16 LDS taps at constant
// offsets from one address per pixel, summed into (FREEZE: subtracted from)
// the wave's sub-block accumulator.
// One pixel of render_box_fixed: its model value and accumulator index.
template <int QF, int PH, int SB, typename T>
__device__ __forceinline__ T render_px_fixed(const float* __restrict__ grid, int base0, int pi, int h, int inv_h,
                                             int r_lo, int c_lo, const T (&wyv)[4], const T (&wxv)[4], T flux,
                                             T bkg, int& idx) {
    const int col = (int)(__umul24((unsigned)pi, (unsigned)inv_h) >> 16);
    const int row = pi - (int)__umul24((unsigned)col, (unsigned)h);
    const int lr = r_lo + row, lc = c_lo + col;
    const float* sp = grid + (base0 + (int)__umul24((unsigned)lc, (unsigned)QF) + lr);
    float gv[16];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) gv[a * 4 + b] = sp[epsf_tap_off(PH, a, b, QF, QF)];
    T rowv[4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
        rowv[a] = wxv[0] * (T)gv[a * 4] + wxv[1] * (T)gv[a * 4 + 1] + wxv[2] * (T)gv[a * 4 + 2] + wxv[3] * (T)gv[a * 4 + 3];
    const T pv = wyv[0] * rowv[0] + wyv[1] * rowv[1] + wyv[2] * rowv[2] + wyv[3] * rowv[3];
    idx = lc * SB + lr;
    return fma(flux, pv, bkg);
}

// The pixels of one star's clipped box (n of them, h rows, local rows /
// columns from r_lo / c_lo) for a star whose whole box lies in its in-range
// window, phase PH and planes of QF x QF cells: the 16 LDS taps at constant
// offsets from one address per pixel, summed into (FREEZE: subtracted from)
// the wave's sub-block accumulator.
template <int QF, int PH, int SB, bool FREEZE, typename T>
__device__ __forceinline__ void render_box_fixed(const float* __restrict__ grid, int base0, int n, int h, int inv_h,
                                                 int r_lo, int c_lo, const T (&wyv)[4], const T (&wxv)[4], T flux,
                                                 T bkg, T* sacc, uint8_t* stouch, int lane) {
    for (int pb = 0; pb < n; pb += 64) {
        const int pi = pb + lane;
        if (pi < n) {
            int idx;
            const T v = render_px_fixed<QF, PH, SB, T>(grid, base0, pi, h, inv_h, r_lo, c_lo, wyv, wxv, flux, bkg, idx);
            if (FREEZE) {
                sacc[idx] -= v;
                stouch[idx] = 1;
            } else {
                sacc[idx] += v;
            }
        }
    }
}

// float renders fit 8 waves / SIMD: two 1024-thread blocks per CU.  QF > 0:
// the ePSF grid is staged with RQF x RQF planes at os 4 (render_box_fixed for
// stars inside their window).
template <typename T, int KIND, int MODE, bool LDSG, int QF = 0, int RTV = RT>
__global__ __launch_bounds__(RBLOCK, sizeof(T) == 4 ? 8 : 4) void k_render(
    ModelConst<T> M, const StarRec<T>* __restrict__ recs, const EpsfRec<T>* __restrict__ ers,
    const CornerRec<T>* __restrict__ crs, const int* __restrict__ tstart, const int* __restrict__ tend,
    const int* __restrict__ list, TileGeo g, int ntiles, const int2* __restrict__ shares,
    const uint8_t* __restrict__ umask, T* __restrict__ dw, const T* __restrict__ w, const T* __restrict__ data,
    T* __restrict__ wt_resid, T* __restrict__ resid, double* __restrict__ partials, Scal* S,
    const int* __restrict__ tmark, int mark_iter) {
    constexpr bool EPSF_LDS = KIND == KIND_EPSF && LDSG;
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ double sred[RWAVE];
    const int ngrid = LDSG ? M.ios_y * M.ios_x * M.Qy * M.Qx : 0;
    float* sgrid = reinterpret_cast<float*>(smem);
    // this block's tiles (its share of the tile-cost scan)
    const int2 share = shares[xcd_remap(blockIdx.x, gridDim.x)];
    const int t0 = share.x, t1 = share.y;
    // the freeze render: a block none of whose tiles a newly frozen star
    // touches stages no grid (the launch follows every accepted step, whether
    // or not a star froze)
    if (MODE == MODE_FREEZE) {
        bool any = false;
        for (int tile = t0 + (int)threadIdx.x; tile < t1; tile += RBLOCK) any |= tmark[tile] == mark_iter;
        if (!__syncthreads_or(any)) return;
    }
    if (LDSG) {
        load_poly(M.poly, sgrid, ngrid, threadIdx.x, RBLOCK);
        __syncthreads();   // the only block barrier: every wave reads the whole grid
    }
    const float* grid = LDSG ? sgrid : M.grids;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    constexpr int SB = RTV / 4, PPL = SB * SB / 64, CPK = 64 / SB;   // pixels per lane, columns per k
    static_assert(SB * SB % 64 == 0, "whole pixels per lane");
    const int sbr = (wv % 4) * SB, sbc = (wv / 4) * SB;
    // this wave's sub-block accumulator (column-major SB x SB) and, for the
    // freeze render, which of its pixels a star touched
    T* sacc = reinterpret_cast<T*>(smem + render_acc_offset<T>(ngrid)) + wv * (SB * SB);
    uint8_t* stouch = smem + render_acc_offset<T>(ngrid) + RWAVE * SB * SB * sizeof(T) + wv * (SB * SB);
    double acc = 0;
    for (int tile = t0; tile < t1; ++tile) {
        // the freeze render: only tiles a newly frozen star touches
        if (MODE == MODE_FREEZE && tmark[tile] != mark_iter) continue;
        const int ty = tile % g.nty, tx = tile / g.nty;
        const int ty0 = ty * RTV + 1, tx0 = tx * RTV + 1;     // tile origin (1-based)
        const int gy0 = ty0 + sbr, gx0 = tx0 + sbc;           // this wave's sub-block origin
        // lane owns pixels (row lane % SB, column CPK k + lane / SB) for the
        // initial and final passes
        const int gy = gy0 + lane % SB;
        bool act[PPL];
        // the pixel inputs of the final pass (data / dw and w) are loaded here,
        // with the union mask rather than after it, so their latency overlaps
        // the star loop instead of following it
        T pin[PPL], pw[PPL];
#pragma unroll
        for (int k = 0; k < PPL; ++k) {
            const int gx = gx0 + CPK * k + lane / SB;
            act[k] = gy <= g.ny && gx <= g.nx;
            const size_t q = (size_t)(gy - 1) + (size_t)(gx - 1) * g.ny;
            pin[k] = pw[k] = T(0);
            if (act[k]) {
                if (MODE == MODE_RESID || MODE == MODE_COST) {
                    pin[k] = dw[q];
                    pw[k] = w[q];
                } else if (MODE == MODE_FINAL) {
                    pin[k] = data[q];
                }
            }
            if ((MODE == MODE_RESID || MODE == MODE_COST) && act[k]) act[k] = umask[q] != 0;
            const int idx = (CPK * k + lane / SB) * SB + lane % SB;
            sacc[idx] = (MODE == MODE_FREEZE && act[k]) ? dw[q] : T(0);
            if (MODE == MODE_FREEZE) stouch[idx] = 0;
        }
        const int s0 = tstart[tile], s1 = tend[tile];
        for (int c0 = s0; c0 < s1; c0 += 64) {
            // each lane inspects one list entry; the wave then takes the stars
            // touching its sub-block in list order.  No block barrier: the 16
            // waves of a tile run independently.
            int ai = 0;
            bool touch = false;
            // fixed-layout ePSF path: the lane's own star's box clipped to the
            // sub-block (r_lo | c_lo << 8 | h << 16 | wd << 24), p / h
            // multiplier, first tap address, phase (-1: general path), flux
            // and stencil weights, taken by the star loop with readlane
            int fbox = 0, finv = 0, fbase = 0, fph = -1;
            T fflux = T(0), fwy[4] = {}, fwx[4] = {};
            if (c0 + lane < s1) {
                ai = list[c0 + lane];
                const StarRec<T> ri = recs[ai];
                const bool use =
                    MODE == MODE_FINAL ? true : (MODE == MODE_FREEZE ? (ri.flags & 2) != 0 : (ri.flags & 1) != 0);
                touch = use && ri.ay - ri.mR <= gy0 + SB - 1 && ri.ay + ri.mR >= gy0 &&
                        ri.ax - ri.mR <= gx0 + SB - 1 && ri.ax + ri.mR >= gx0;
                if constexpr (EPSF_LDS && QF > 0) if (touch) {
                    const EpsfRec<T>& er = ers[ai];
                    const int r_lo = max(ri.ay - ri.mR - gy0, 0), r_hi = min(ri.ay + ri.mR - gy0, SB - 1);
                    const int c_lo = max(ri.ax - ri.mR - gx0, 0), c_hi = min(ri.ax + ri.mR - gx0, SB - 1);
                    const int dy0 = gy0 - ri.ay, dx0 = gx0 - ri.ax;
                    const int h = r_hi - r_lo + 1;
                    // the whole clipped box inside the star's in-range window
                    const bool inw = er.ok && dy0 + r_lo >= er.dylo && dy0 + r_hi <= er.dyhi &&
                                     dx0 + c_lo >= er.dxlo && dx0 + c_hi <= er.dxhi;
                    if (inw) {
                        fbox = r_lo | (c_lo << 8) | (h << 16) | ((c_hi - c_lo + 1) << 24);
                        finv = RENDER_INV_H[h];
                        fbase = er.base + dy0 + dx0 * QF;
                        fph = er.phase;
                        fflux = ri.flux;
#pragma unroll
                        for (int t = 0; t < 4; ++t) {
                            fwy[t] = er.wy[t];
                            fwx[t] = er.wx[t];
                        }
                    }
                }
            }
            unsigned long long mine = __ballot(touch);
            while (mine) {
                const int j = __builtin_amdgcn_readfirstlane(__builtin_ctzll(mine));
                mine &= mine - 1;
                if constexpr (EPSF_LDS && QF > 0) {
                    const int ph = __builtin_amdgcn_readlane(fph, j);
                    if (ph >= 0) {
                        const int bx = __builtin_amdgcn_readlane(fbox, j);
                        const int r_lo = bx & 255, c_lo = (bx >> 8) & 255, h = (bx >> 16) & 255, wd = bx >> 24;
                        const int inv_h = __builtin_amdgcn_readlane(finv, j);
                        const int base0 = __builtin_amdgcn_readlane(fbase, j);
                        const T flux = lane_value(fflux, j);
                        T wyv[4], wxv[4];
#pragma unroll
                        for (int t = 0; t < 4; ++t) {
                            wyv[t] = lane_value(fwy[t], j);
                            wxv[t] = lane_value(fwx[t], j);
                        }
#define CP_RBF(PHV)                                                                                          \
    case PHV:                                                                                                \
        render_box_fixed<QF, PHV, SB, MODE == MODE_FREEZE, T>(grid, base0, h * wd, h, inv_h, r_lo, c_lo, wyv, wxv, \
                                                          flux, M.bkg, sacc, stouch, lane);               \
        break;
                        switch (ph) {
                            CP_RBF(0) CP_RBF(1) CP_RBF(2) CP_RBF(3) CP_RBF(4) CP_RBF(5) CP_RBF(6) CP_RBF(7)
                            CP_RBF(8) CP_RBF(9) CP_RBF(10) CP_RBF(11) CP_RBF(12) CP_RBF(13) CP_RBF(14) CP_RBF(15)
                        }
#undef CP_RBF
                        continue;
                    }
                }
                const int a = __builtin_amdgcn_readlane(ai, j);
                const StarRec<T> r = recs[a];
                // the model box clipped to the sub-block, in local rows / columns;
                // its h x wd pixels are spread over the lanes (pixel p: column
                // p / h, row p % h), so no lane idles outside the box
                const int r_lo = max(r.ay - r.mR - gy0, 0), r_hi = min(r.ay + r.mR - gy0, SB - 1);
                const int c_lo = max(r.ax - r.mR - gx0, 0), c_hi = min(r.ax + r.mR - gx0, SB - 1);
                const int h = r_hi - r_lo + 1, n = h * (c_hi - c_lo + 1);
                const int inv_h = RENDER_INV_H[h];   // p / h == (p inv_h) >> 16 for p < 256
                // this star's ePSF stencil (epsf_eval with the record in registers)
                int toff[16], wlo[2], whi[2];
                bool sok = false;
                T wyv[4], wxv[4];
                if constexpr (EPSF_LDS) {
                    const EpsfRec<T>& er = ers[a];
                    sok = er.ok;
                    wlo[0] = er.dylo;
                    whi[0] = er.dyhi;
                    wlo[1] = er.dxlo;
                    whi[1] = er.dxhi;
#pragma unroll
                    for (int ta = 0; ta < 4; ++ta) {
                        wyv[ta] = er.wy[ta];
                        wxv[ta] = er.wx[ta];
#pragma unroll
                        for (int tb = 0; tb < 4; ++tb) toff[ta * 4 + tb] = er.ra[ta] + er.ca[tb];
                    }
                }
                for (int base = 0; base < n; base += 64) {
                    const int pi = base + lane;
                    if (pi < n) {
                        const int col = (pi * inv_h) >> 16, row = pi - col * h;
                        const int lr = r_lo + row, lc = c_lo + col;
                        const int dy = gy0 + lr - r.ay, dx = gx0 + lc - r.ax;
                        T v;
                        if constexpr (EPSF_LDS) {
                            // all 16 LDS taps issued before the first is used
                            const bool inb = sok && dy >= wlo[0] && dy <= whi[0] && dx >= wlo[1] && dx <= whi[1];
                            const int dyc = min(max(dy, wlo[0]), whi[0]), dxc = min(max(dx, wlo[1]), whi[1]);
                            const float* sp = grid + (dyc + dxc * M.Qy);
                            float gv[16];
#pragma unroll
                            for (int t = 0; t < 16; ++t) gv[t] = sp[toff[t]];
                            // schedule: the 16 LDS reads back to back, then the arithmetic
                            __builtin_amdgcn_sched_group_barrier(0x0100, 16, 0);
                            __builtin_amdgcn_sched_group_barrier(0x0002, 40, 0);
                            T rowv[4];
#pragma unroll
                            for (int ta = 0; ta < 4; ++ta)
                                rowv[ta] = wxv[0] * (T)gv[ta * 4] + wxv[1] * (T)gv[ta * 4 + 1] + wxv[2] * (T)gv[ta * 4 + 2] +
                                           wxv[3] * (T)gv[ta * 4 + 3];
                            const T pv4 = wyv[0] * rowv[0] + wyv[1] * rowv[1] + wyv[2] * rowv[2] + wyv[3] * rowv[3];
                            const T pv = inb ? pv4 : M.fill;
                            v = fma(r.flux, pv, M.bkg);
                        } else {
                            v = star_value<T, KIND>(M, r, T(dy) - r.fy, T(dx) - r.fx, grid, crs + a);
                        }
                        const int idx = lc * SB + lr;
                        if (MODE == MODE_FREEZE) {
                            sacc[idx] -= v;
                            stouch[idx] = 1;
                        } else {
                            sacc[idx] += v;
                        }
                    }
                }
            }
        }
#pragma unroll
        for (int k = 0; k < PPL; ++k) {
            if (!act[k]) continue;
            const int gx = gx0 + CPK * k + lane / SB;
            const size_t q = (size_t)(gy - 1) + (size_t)(gx - 1) * g.ny;
            const int idx = (CPK * k + lane / SB) * SB + lane % SB;
            const T val = sacc[idx];
            if (MODE == MODE_RESID || MODE == MODE_COST) {
                T r = val - pin[k];
                T wr = sqrt(pw[k]) * r;
                wt_resid[q] = wr;
                acc += (double)wr * (double)wr;
            } else if (MODE == MODE_FREEZE) {
                if (stouch[idx]) dw[q] = val;
            } else {
                resid[q] = pin[k] - val;
            }
        }
    }
    if (MODE == MODE_RESID || MODE == MODE_COST) block_store<RWAVE>(acc, partials, sred);
}

// ---------------------------------------------------------------------------
// fill + J^T r (one wave per star; _fill_stamps! :342-543, apply_JT! :55-81,
// g-cosine :1195-1206 and _g_tol_local :789-801)
// ---------------------------------------------------------------------------
// GL lanes per star (16 or 64), EPL stamp entries per lane.
template <int GL>
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
    for (int o = GL / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// one wave per star; float fills with one or two stamp entries per lane fit
// 64 VGPRs (8 waves / SIMD)
template <typename T, int KIND, bool FINAL, bool LDSG, int EPL, int GL, typename VT>
__global__ __launch_bounds__(FBLOCK, sizeof(T) == 4 && EPL <= 2 ? 8 : 4) void k_fill(
    ModelConst<T> M, const StarRec<T>* __restrict__ recs, const EpsfRec<T>* __restrict__ ers,
    const CornerRec<T>* __restrict__ crs, const int* __restrict__ lst, int na, int R, int S, FreeMap fm,
    const T* __restrict__ w, int ny, int nx, const T* __restrict__ wt_resid, VT* __restrict__ V,
    double* __restrict__ colnorm, double* __restrict__ b_scaled, uint8_t* __restrict__ gloc, double g_tol,
    double* __restrict__ partials, Scal* Sc) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ double sred[FWAVE];
    float* sgrid = reinterpret_cast<float*>(smem);
    if (LDSG) {
        load_poly(M.poly, sgrid, M.ios_y * M.ios_x * M.Qy * M.Qx, threadIdx.x, FBLOCK);
        __syncthreads();
    }
    const float* grid = LDSG ? sgrid : M.grids;
    constexpr int GPB = FBLOCK / GL;   // stars in flight per block
    const int sub = threadIdx.x % GL;
    const int S2 = S * S;
    const int p = fm.p;
    const double C_r = FINAL ? 0.0 : Sc->cost / Sc->dof;
    double bsq = 0, gmax = 0;
    // stars: the list lst (Sc->n_jtu_stars entries: the live stars of the last
    // compaction; stars outside it are frozen and keep their columns), or all
    const int cnt = lst ? Sc->n_jtu_stars : na;
    for (int li = xcd_remap(blockIdx.x, gridDim.x) * GPB + threadIdx.x / GL; li < cnt; li += gridDim.x * GPB) {
        const int a = lst ? lst[li] : li;
        const StarRec<T> r = recs[a];
        if (!FINAL && !(r.flags & 1)) {   // uniform within the group
            if (sub < VP) b_scaled[(size_t)a * VP + sub] = 0.0;
            continue;
        }
        EpsfRec<T> er;
        if (KIND == KIND_EPSF && LDSG) er = ers[a];
        // raw weighted gradient columns; once the column norms are known, b =
        // J^T r is summed from the stored (storage-precision) values of J, so
        // b and the normal matrix the inner solve forms from V describe the
        // same J
        T raw[3][EPL];
        double wrs[EPL];
        double cn2[3] = {0, 0, 0};
#pragma unroll
        for (int e = 0; e < EPL; ++e) {
            const int m = sub + GL * e;
            raw[0][e] = raw[1][e] = raw[2][e] = T(0);
            wrs[e] = 0;
            const int dy = (m % S) - R, dx = (m / S) - R;
            const int gy = r.ay + dy, gx = r.ax + dx;
            const bool on = m < S2 && gy >= 1 && gy <= ny && gx >= 1 && gx <= nx;
            const size_t q = on ? (size_t)(gy - 1) + (size_t)(gx - 1) * ny : 0;
            const T wv = on ? w[q] : T(0);
            if (!(wv > T(0))) continue;
            T g3[3];
            if constexpr (KIND == KIND_EPSF && LDSG) {
                T pv, pdv, pdu;
                epsf_eval<T, true>(M, er, grid, dy, dx, pv, pdv, pdu);
                g3[0] = r.flux * (-(M.osy * pdv));
                g3[1] = r.flux * (-(M.osx * pdu));
                g3[2] = pv;
            } else {
                star_grad<T, KIND>(M, r, T(dy) - r.fy, T(dx) - r.fx, grid, crs + a, g3[0], g3[1], g3[2]);
            }
            const T sw = sqrt(wv);
            wrs[e] = FINAL ? 0.0 : (double)wt_resid[q];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                if (k < p) {
                    int c = gc_of(fm, k);
                    T gv = c == 1 ? g3[0] : (c == 2 ? g3[1] : g3[2]);
                    T v = sw * gv;
                    raw[k][e] = v;
                    cn2[k] += (double)v * (double)v;
                }
            }
        }
        double cn[3] = {1, 1, 1}, bk[3] = {0, 0, 0}, inv[3] = {0, 0, 0};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            if (k < p) {
                double c = sqrt(group_sum<GL>(cn2[k]));
                cn[k] = c < DEPS ? DEPS : c;
                inv[k] = 1.0 / cn[k];
            }
        }
        vec4_t<VT>* Vrow = reinterpret_cast<vec4_t<VT>*>(V) + (size_t)a * S2;
        double braw[3] = {0, 0, 0};
#pragma unroll
        for (int e = 0; e < EPL; ++e) {
            const int m = sub + GL * e;
            if (m < S2) {
                vec4_t<VT> v;
                v.x = (VT)((double)raw[0][e] * inv[0]);
                v.y = (VT)((double)raw[1][e] * inv[1]);
                v.z = (VT)((double)raw[2][e] * inv[2]);
                v.w = VT(0);
                Vrow[m] = v;
                braw[0] += (double)v.x * wrs[e];
                braw[1] += (double)v.y * wrs[e];
                braw[2] += (double)v.z * wrs[e];
            }
        }
        if (!FINAL) {
#pragma unroll
            for (int k = 0; k < 3; ++k)
                if (k < p) bk[k] = group_sum<GL>(braw[k]);
        }
        if (sub == 0) {
            double gn = 0;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                if (k < p) {
                    colnorm[(size_t)a * p + k] = cn[k];
                    if (!FINAL) {
                        b_scaled[(size_t)a * VP + k] = bk[k];
                        bsq += bk[k] * bk[k];
                        double bt = cn[k] * bk[k];
                        gn = fmax(gn, fabs(bt) / (sqrt(cn[k] * cn[k] * C_r) + DEPS));
                    }
                }
            }
            if (!FINAL) {
                gloc[a] = gn <= g_tol;
                gmax = fmax(gmax, gn);
            }
        }
    }
    if (!FINAL) {
        // partials[0, nb): |b|^2 per block, partials[nb, 2 nb): max g-cosine per block
        block_store<FWAVE>(bsq, partials, sred);
        const double m = block_max<FWAVE>(gmax, sred);
        if (threadIdx.x == 0) partials[gridDim.x + blockIdx.x] = m;
    }
}

// ---------------------------------------------------------------------------
// inner solver: LSQR (Paige & Saunders 1982) on the stamp operator
// ---------------------------------------------------------------------------
// LSQR scalar recurrences of iteration `it` (one thread), from tt = |J^T u -
// beta v|^2 and ww = |w|^2 of that iteration (Paige & Saunders 1982, with the
// Krylov.jl-style stopping tests).
__device__ void lsqr_advance(Scal* Sc, double tt, double ww, int it) {
#pragma clang fp contract(off)   // rounded alike wherever it is inlined
    const double alpha = Sc->alpha, beta = Sc->beta_next;
    const double lam = Sc->lam;
    double alpha_new = alpha;
    if (beta != 0) {
        Sc->anorm2 += alpha * alpha + beta * beta + (lam > 0 ? lam * lam : 0.0);
        alpha_new = sqrt(tt);
    }
    double c1, s1, rhobar1;
    sym_givens(Sc->rhobar, lam, c1, s1, rhobar1);
    double psi = s1 * Sc->phibar;
    double phibar = c1 * Sc->phibar;
    double c, s, rho;
    sym_givens(rhobar1, beta, c, s, rho);
    double theta = s * alpha_new;
    Sc->rhobar = -c * alpha_new;
    double phi = c * phibar;
    phibar = s * phibar;
    Sc->phibar = phibar;
    Sc->xenorm2 += phi * phi;
    Sc->errv[it % 5] = phi;
    double err_lbnd = 0;
    if (it >= 5) {
        double e2 = 0;
        for (int i = 0; i < 5; ++i) e2 += Sc->errv[i] * Sc->errv[i];
        err_lbnd = sqrt(e2);
    }
    double tau_r = s * phi;
    Sc->dnorm2 += ww / (rho * rho);
    Sc->sig = phi / rho;
    Sc->tau = theta / rho;
    double delta = Sc->s2 * rho, gambar = -Sc->c2 * rho, rhs = phi - delta * Sc->z;
    double zbar = rhs / gambar;
    double xnorm = sqrt(Sc->xnorm2 + zbar * zbar);
    double c2, s2, gamma;
    sym_givens(gambar, theta, c2, s2, gamma);
    Sc->c2 = c2;
    Sc->s2 = s2;
    Sc->z = rhs / gamma;
    Sc->xnorm2 += Sc->z * Sc->z;
    double anorm = sqrt(Sc->anorm2);
    double acond = anorm * sqrt(Sc->dnorm2);
    Sc->res2 += psi * psi;
    double rnorm = sqrt(phibar * phibar + Sc->res2);
    double arnorm = alpha_new * fabs(tau_r);
    double beta1 = Sc->beta1;
    double test1 = rnorm / beta1, test2 = arnorm / (anorm * rnorm), test3 = 1 / acond;
    double t1 = test1 / (1 + anorm * xnorm / beta1);
    const double axtol = 1.4901161193847656e-08, etol = axtol, ctol = axtol;
    double rtol = Sc->btol + axtol * anorm * xnorm / beta1;
    bool ill = (1 + test3 <= 1) || (test3 <= ctol);
    bool solved = (1 + test2 <= 1) || (test2 <= axtol) || (arnorm <= Sc->atol) || (1 + t1 <= 1) ||
                  (test1 <= rtol) || (it >= 5 && err_lbnd <= etol * sqrt(Sc->xenorm2));
    Sc->alpha = alpha_new;
    Sc->beta_cur = beta;
    Sc->iter = it;
    Sc->done = solved || ill || it >= Sc->itmax;
}

// LSMR scalar recurrences of iteration `it` (one thread), Fong & Saunders
// 2011 with the Krylov.jl stopping tests, from tt = |t|^2 (t = alpha_new
// v_new) and the products q[0..5] = |x|^2, x.h, x.hbar, |h|^2, h.hbar,
// |hbar|^2 of the vectors after the previous iteration's update, which give
// |x_it| once this iteration's del and sigma are known.
__device__ void lsmr_advance(Scal* Sc, double tt, const double* q, int it) {
#pragma clang fp contract(off)   // rounded alike wherever it is inlined
    const double beta = Sc->beta_next, lam = Sc->lam;
    double alpha = Sc->alpha;
    if (beta != 0) alpha = sqrt(tt);
    double chat, shat, alphahat;
    sym_givens(Sc->alphabar, lam, chat, shat, alphahat);
    const double rhoold = Sc->rho;
    double c, s, rho;
    sym_givens(alphahat, beta, c, s, rho);
    const double thetanew = s * alpha;
    Sc->alphabar = c * alpha;
    const double rhobarold = Sc->rhobar_m, zetaold = Sc->zeta;
    const double thetabar = Sc->sbar * rho, rhotemp = Sc->cbar * rho;
    double cbar, sbar, rhobar;
    sym_givens(rhotemp, thetanew, cbar, sbar, rhobar);
    Sc->cbar = cbar;
    Sc->sbar = sbar;
    Sc->rhobar_m = rhobar;
    Sc->rho = rho;
    const double zeta = cbar * Sc->zetabar;
    Sc->zeta = zeta;
    Sc->zetabar = -sbar * Sc->zetabar;
    Sc->xenorm2 += zeta * zeta;
    Sc->errv[it % 5] = zeta;
    double err_lbnd = 0;
    if (it >= 5) {
        double e2 = 0;
        for (int i = 0; i < 5; ++i) e2 += Sc->errv[i] * Sc->errv[i];
        err_lbnd = sqrt(e2);
    }
    // pending update: hbar <- h - del hbar, x <- x + sig hbar, h <- v - tau h
    const double del = thetabar * rho / (rhoold * rhobarold);
    const double sig = zeta / (rho * rhobar);
    Sc->del = del;
    Sc->sig = sig;
    Sc->tau = thetanew / rho;
    const double hb2 = q[3] - 2 * del * q[4] + del * del * q[5];
    const double xhb = q[1] - del * q[2];
    const double xnorm = sqrt(fmax(0.0, q[0] + 2 * sig * xhb + sig * sig * hb2));
    // ||r|| estimate
    const double betaacute = chat * Sc->betadd, betacheck = -shat * Sc->betadd;
    const double betahat = c * betaacute;
    Sc->betadd = -s * betaacute;
    const double thetatildeold = Sc->thetatilde;
    double ctildeold, stildeold, rhotildeold;
    sym_givens(Sc->rhodold, thetabar, ctildeold, stildeold, rhotildeold);
    Sc->thetatilde = stildeold * rhobar;
    Sc->rhodold = ctildeold * rhobar;
    Sc->betad = -stildeold * Sc->betad + ctildeold * betahat;
    Sc->tautildeold = (zetaold - thetatildeold * Sc->tautildeold) / rhotildeold;
    const double taud = (zeta - Sc->thetatilde * Sc->tautildeold) / Sc->rhodold;
    Sc->dsum += betacheck * betacheck;
    const double bd = Sc->betad - taud;
    const double rnorm = sqrt(Sc->dsum + bd * bd + Sc->betadd * Sc->betadd);
    // ||A|| and cond(A)
    Sc->anorm2 += beta * beta;
    const double anorm = sqrt(Sc->anorm2);
    Sc->anorm2 += alpha * alpha;
    Sc->maxrbar = fmax(Sc->maxrbar, rhobarold);
    if (it > 1) Sc->minrbar = fmin(Sc->minrbar, rhobarold);
    const double acond = fmax(Sc->maxrbar, rhotemp) / fmin(Sc->minrbar, rhotemp);
    const double arnorm = fabs(Sc->zetabar);
    const double beta1 = Sc->beta1;
    const double test1 = rnorm / beta1, test2 = arnorm / (anorm * rnorm), test3 = 1 / acond;
    const double t1 = test1 / (1 + anorm * xnorm / beta1);
    const double axtol = 1.4901161193847656e-08, etol = axtol, ctol = axtol;
    const double rtol = Sc->btol + axtol * anorm * xnorm / beta1;
    const bool ill = (1 + test3 <= 1) || (test3 <= ctol);
    const bool solved = (1 + test2 <= 1) || (test2 <= axtol) || (arnorm <= Sc->atol) || (1 + t1 <= 1) ||
                        (test1 <= rtol) || (it >= 5 && err_lbnd <= etol * sqrt(Sc->xenorm2));
    Sc->alpha = alpha;
    Sc->beta_cur = beta;
    Sc->iter = it;
    Sc->done = solved || ill || it >= Sc->itmax;
}

// ---------------------------------------------------------------------------
// k_finish: one block sums the per-block partials of the preceding launch in
// block order and updates the device scalars
// ---------------------------------------------------------------------------
enum : int {
    FIN_COST = 0, FIN_COST_CAND = 1, FIN_FILL = 2, FIN_ACCEPT = 5, FIN_COUNT = 6, FIN_CGQ = 7, FIN_CGS = 8
};
constexpr int FIN_BLOCK = 1024;

// Block-order sums of the partials of a launch of nb blocks by one
// FIN_BLOCK-thread block: ta (and tb, a max for the fill's g-cosine) in
// thread 0.
template <int OP>
__device__ __forceinline__ void fin_reduce(const double* __restrict__ part, const double* __restrict__ part2, int nb,
                                           double (*sh)[FIN_BLOCK / 64], double& ta, double& tb) {
    double a = 0, b = 0;
    for (int i = threadIdx.x; i < nb; i += FIN_BLOCK) {
        a += part[i];
        if (OP == FIN_FILL) b = fmax(b, part[nb + i]);
        if (OP == FIN_ACCEPT) b += part[nb + i];
        if (OP == FIN_CGS || OP == FIN_CGQ) b += part2[i];
    }
    a = wave_sum(a);
    b = OP == FIN_FILL ? wave_max(b) : wave_sum(b);
    if ((threadIdx.x & 63) == 0) {
        sh[0][threadIdx.x >> 6] = a;
        sh[1][threadIdx.x >> 6] = b;
    }
    __syncthreads();
    ta = 0;
    tb = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < FIN_BLOCK / 64; ++i) {
            ta += sh[0][i];
            tb = OP == FIN_FILL ? fmax(tb, sh[1][i]) : tb + sh[1][i];
        }
}

// mirror (optional): the updated scalars also go to this pinned host record.
template <int OP>
static __global__ __launch_bounds__(FIN_BLOCK) void k_finish(Scal* Sc, const double* __restrict__ part,
                                                             const double* __restrict__ part2, int nb, int it,
                                                             Scal* mirror) {
    if ((OP == FIN_CGQ || OP == FIN_CGS) && Sc->done) return;
    __shared__ double sh[2][FIN_BLOCK / 64];
    double ta, tb;
    fin_reduce<OP>(part, part2, nb, sh, ta, tb);
    if (threadIdx.x == 0) {
        if (OP == FIN_COST) Sc->cost = ta;
        if (OP == FIN_COST_CAND) Sc->cost_cand = ta;
        if (OP == FIN_FILL) {
            Sc->bnorm2 = ta;
            Sc->gmax_bits = (unsigned long long)__double_as_longlong(tb);
        }
        if (OP == FIN_CGQ) {
            // ta = p . H p, tb = |p|^2: the step a = gamma / (p'Hp + lam^2 |p|^2)
            const double delta = ta + Sc->lam * Sc->lam * tb;
            if (!(delta > 0)) {
                Sc->done = 1;
            } else {
                Sc->cg_alpha = Sc->gamma / delta;
            }
        }
        if (OP == FIN_CGS) {
            // ta = |s_new|^2, tb = s_new . p: the next direction p + = s_new + beta p
            const double beta = ta / Sc->gamma;
            Sc->pnorm2 = ta + 2 * beta * tb + beta * beta * Sc->pnorm2;
            Sc->cg_beta = beta;
            Sc->gamma = ta;
            Sc->iter = it;
            Sc->done = sqrt(ta) <= Sc->atol * Sc->s0norm || it >= Sc->itmax;
        }
        if (OP == FIN_ACCEPT) {
            // the live and newly frozen counts; the accepted candidate's cost
            // becomes the current cost (its weighted residual is reused as the
            // next iteration's residual)
            Sc->n_live = (int)ta;
            Sc->n_frozen_now = (int)tb;
            Sc->cost = Sc->cost_cand;
        }
        if (OP == FIN_COUNT) Sc->n_union = (unsigned long long)ta;
    }
    if (mirror) {   // uniform
        __syncthreads();
        mirror_scal(Sc, mirror);
    }
}

// fpart (the first trial of an outer iteration): the fill's partials over fnb
// blocks are summed first (|b|^2, max g-cosine: the FIN_FILL reduction).
static __global__ __launch_bounds__(FIN_BLOCK) void k_lsqr_init(Scal* S, double lam, int itmax, double atol,
                                                                double btol, int solver,
                                                                const double* __restrict__ fpart, int fnb) {
    if (fpart) {   // uniform
        __shared__ double sh[2][FIN_BLOCK / 64];
        double ta, tb;
        fin_reduce<FIN_FILL>(fpart, nullptr, fnb, sh, ta, tb);
        if (threadIdx.x == 0) {
            S->bnorm2 = ta;
            S->gmax_bits = (unsigned long long)__double_as_longlong(tb);
        }
    }
    if (threadIdx.x != 0) return;
    S->lam = lam;
    S->itmax = itmax;
    S->atol = atol;
    S->btol = btol;
    S->solver = solver;
    S->sig = S->tau = S->del = 0;
    S->iter = 0;
    S->xnorm2 = 0;
    for (int i = 0; i < 5; ++i) S->errv[i] = 0;
    S->xenorm2 = 0;
    if (solver == CP_SOLVER_CGLS) {
        // s_0 = J' b - lam^2 x_0 = -b_scaled, p_0 = s_0
        S->gamma = S->bnorm2;
        S->pnorm2 = S->bnorm2;
        S->s0norm = sqrt(S->bnorm2);
        S->cg_beta = 0;
        S->cg_alpha = 0;
        S->done = !(S->bnorm2 > 0);
        return;
    }
    double beta1 = sqrt(S->cost);
    double alpha = beta1 > 0 ? sqrt(S->bnorm2) / beta1 : 0.0;
    S->beta1 = beta1;
    S->alpha = alpha;
    S->beta_cur = beta1;
    S->beta_next = 0;
    S->anorm2 = alpha * alpha;
    S->done = (beta1 == 0 || alpha == 0);
    if (solver == CP_SOLVER_LSMR) {
        S->zetabar = alpha * beta1;
        S->alphabar = alpha;
        S->rho = 1;
        S->rhobar_m = 1;
        S->cbar = 1;
        S->sbar = 0;
        S->betadd = beta1;
        S->betad = 0;
        S->rhodold = 1;
        S->tautildeold = 0;
        S->thetatilde = 0;
        S->zeta = 0;
        S->dsum = 0;
        S->maxrbar = 0;
        S->minrbar = 1e100;
        return;
    }
    S->rhobar = alpha;
    S->phibar = beta1;
    S->dnorm2 = S->z = S->res2 = 0;
    S->c2 = -1;
    S->s2 = 0;
}

// One star's final x update, capped step and candidate theta
// (psf_photometry_simultaneous.jl:1224-1230, _cap_position_step! :700-716), from
// the solve's final scalars S.  The last iteration's pending update is
// x + sig (w - del hbar) (LSQR: del = 0; CG: sig = 0, x is already final).  No
// fused multiply-adds, so every kernel that inlines it rounds alike.
__device__ __forceinline__ void step_star(int a, int p, int k_y, int k_x, double max_step,
                                          const double* __restrict__ wv, const double* __restrict__ hb,
                                          double* __restrict__ x, const double* __restrict__ colnorm,
                                          const double* __restrict__ theta, double* __restrict__ dsc,
                                          double* __restrict__ theta_c, const uint8_t* __restrict__ live,
                                          const Scal* S) {
#pragma clang fp contract(off)
    const size_t base = (size_t)a * p;
    const bool any_iter = S->iter > 0;
    const double sig = S->sig, del = S->del;
    double d[3];
    for (int k = 0; k < p; ++k) {
        const double h = hb ? wv[base + k] - del * hb[base + k] : wv[base + k];
        double xv = (any_iter && live[a]) ? x[base + k] + sig * h : 0.0;
        x[base + k] = xv;
        d[k] = xv / colnorm[base + k];
    }
    if (k_y >= 0 || k_x >= 0) {
        double dy = k_y >= 0 ? d[k_y] : 0.0, dx = k_x >= 0 ? d[k_x] : 0.0;
        double len = sqrt(dy * dy + dx * dx);
        if (len > max_step) {
            double sc = max_step / len;
            if (k_y >= 0) d[k_y] *= sc;
            if (k_x >= 0) d[k_x] *= sc;
        }
    }
    for (int k = 0; k < p; ++k) {
        dsc[base + k] = colnorm[base + k] * d[k];
        theta_c[base + k] = theta[base + k] + d[k];
    }
}

static __global__ void k_step(int na, int p, int k_y, int k_x, double max_step, const double* __restrict__ wv,
                       const double* __restrict__ hb, double* __restrict__ x, const double* __restrict__ colnorm,
                       const double* __restrict__ theta, double* __restrict__ dsc, double* __restrict__ theta_c,
                       const uint8_t* __restrict__ live, const Scal* Sc) {
    const int a = blockIdx.x * BLOCK + threadIdx.x;
    if (a >= na) return;
    step_star(a, p, k_y, k_x, max_step, wv, hb, x, colnorm, theta, dsc, theta_c, live, Sc);
}

// Accepted step: theta, D, per-star freeze tests (:1246-1283, _x_tol_local :810-823).
static __global__ void k_accept(int na, int p, int iter, double x_tol, double* __restrict__ theta,
                         const double* __restrict__ theta_c, double* __restrict__ D,
                         const double* __restrict__ colnorm, const double* __restrict__ dsc,
                         const uint8_t* __restrict__ gloc, uint8_t* __restrict__ live,
                         int* __restrict__ frozen_at, uint8_t* __restrict__ conv, double* __restrict__ partials) {
    int a = blockIdx.x * BLOCK + threadIdx.x;
    int nl = 0, nf = 0;
    if (a < na) {
        const size_t base = (size_t)a * p;
        for (int k = 0; k < p; ++k) {
            theta[base + k] = theta_c[base + k];
            D[base + k] = fmax(D[base + k], colnorm[base + k]);
        }
        if (live[a]) {
            double num = 0, xn = 0;
            for (int k = 0; k < p; ++k) {
                double v = (D[base + k] / colnorm[base + k]) * dsc[base + k];
                num += v * v;
                double xv = D[base + k] * theta[base + k];
                xn += xv * xv;
            }
            bool xt = sqrt(num) <= x_tol * (sqrt(xn) + x_tol);
            if (xt || gloc[a]) {
                live[a] = 0;
                frozen_at[a] = iter;
                conv[a] = 1;
                nf = 1;
            } else {
                nl = 1;
            }
        }
    }
    __shared__ double sred[NWAVE];
    // partials[0, nb): live count, partials[nb, 2 nb): newly frozen count
    block_store((double)nl, partials, sred);
    const double f = block_sum((double)nf, sred);
    if (threadIdx.x == 0) partials[gridDim.x + blockIdx.x] = f;
}

// _freeze_remaining! (:767-776).
static __global__ void k_freeze_all(int na, int iter, uint8_t* __restrict__ live, int* __restrict__ frozen_at,
                             uint8_t* __restrict__ conv) {
    int a = blockIdx.x * BLOCK + threadIdx.x;
    if (a >= na || !live[a]) return;
    live[a] = 0;
    frozen_at[a] = iter;
    conv[a] = 1;
}

// Per-star error block, covariance and errors (:1305-1329, covariance!
// levenberg_marquardt.jl:366-388).  cov_mode 1 known, 2 reweighted, -1 none.
// ERR_GL lanes per star stride over its stamp rows (consecutive rows, so a
// group's loads are contiguous) and sum the block's distinct entries over the
// group; every lane of the group then runs the same small inverse and the
// group's first lane writes.
constexpr int ERR_GL = 16;
template <typename T, int P, typename VT>
static __global__ __launch_bounds__(BLOCK) void k_errors(int na, int R, int S, const int* __restrict__ ay,
                                                         const int* __restrict__ ax, const T* __restrict__ w,
                                                         int ny, int nx, const VT* __restrict__ V,
                                                         const double* __restrict__ colnorm, int cov_mode,
                                                         double fac_rw, double* __restrict__ errs) {
    const int a = blockIdx.x * (BLOCK / ERR_GL) + threadIdx.x / ERR_GL;
    const int sub = threadIdx.x % ERR_GL;
    if (a >= na) return;   // uniform over the group: the group sums stay inside it
    constexpr int p = P;
    const int S2 = S * S;
    double blk[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const int ya = ay[a], xa = ax[a];
    for (int m = sub; m < S2; m += ERR_GL) {
        const int dx = m / S;
        const int gy = ya + (m - dx * S) - R, gx = xa + dx - R;
        if (gy < 1 || gy > ny || gx < 1 || gx > nx) continue;
        if (!(w[(size_t)(gy - 1) + (size_t)(gx - 1) * ny] > T(0))) continue;
        const VT* Vm = V + ((size_t)a * S2 + m) * VP;
#pragma unroll
        for (int k = 0; k < p; ++k)
#pragma unroll
            for (int l = k; l < p; ++l) blk[k * 3 + l] += (double)Vm[k] * (double)Vm[l];
    }
#pragma unroll
    for (int k = 0; k < p; ++k)
#pragma unroll
        for (int l = k; l < p; ++l) {
            blk[k * 3 + l] = group_sum<ERR_GL>(blk[k * 3 + l]);
            blk[l * 3 + k] = blk[k * 3 + l];
        }
    const double* cn = colnorm + (size_t)a * p;
    double tr = 0;
#pragma unroll
    for (int k = 0; k < p; ++k)
#pragma unroll
        for (int l = 0; l < p; ++l) blk[k * 3 + l] *= cn[k] * cn[l];
#pragma unroll
    for (int k = 0; k < p; ++k) tr += blk[k * 3 + k];
#pragma unroll
    for (int k = 0; k < p; ++k) blk[k * 3 + k] += 1e-12 * tr;
    double cov[9];
    // Cholesky inverse, else eigen pseudo-inverse (pinv)
    double L[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    bool pd = true;
#pragma unroll
    for (int j = 0; j < p; ++j) {
        double s = blk[j * 3 + j];
#pragma unroll
        for (int k = 0; k < j; ++k) s -= L[j * 3 + k] * L[j * 3 + k];
        pd = pd && (s > 0);
        const double d = pd ? sqrt(s) : 1.0;
        L[j * 3 + j] = d;
#pragma unroll
        for (int i = j + 1; i < p; ++i) {
            double t = blk[j * 3 + i];
#pragma unroll
            for (int k = 0; k < j; ++k) t -= L[i * 3 + k] * L[j * 3 + k];
            L[i * 3 + j] = t / d;
        }
    }
    if (pd) {
#pragma unroll
        for (int j = 0; j < p; ++j) {
            double y[3], xs[3];
#pragma unroll
            for (int i = 0; i < p; ++i) {
                double s = (i == j);
                for (int k = 0; k < i; ++k) s -= L[i * 3 + k] * y[k];
                y[i] = s / L[i * 3 + i];
            }
#pragma unroll
            for (int i = p - 1; i >= 0; --i) {
                double s = y[i];
#pragma unroll
                for (int k = i + 1; k < p; ++k) s -= L[k * 3 + i] * xs[k];
                xs[i] = s / L[i * 3 + i];
            }
#pragma unroll
            for (int i = 0; i < p; ++i) cov[i * 3 + j] = xs[i];
        }
    } else {
        double A[9], E[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int i = 0; i < 9; ++i) A[i] = blk[i];
#pragma unroll
        for (int i = 0; i < p; ++i) E[i * 3 + i] = 1;
        for (int sweep = 0; sweep < 60; ++sweep) {
            double off = 0;
#pragma unroll
            for (int i = 0; i < p; ++i)
#pragma unroll
                for (int j = i + 1; j < p; ++j) off += A[i * 3 + j] * A[i * 3 + j];
            if (off < 1e-300) break;
#pragma unroll
            for (int i = 0; i < p; ++i)
#pragma unroll
                for (int j = i + 1; j < p; ++j) {
                    double aij = A[i * 3 + j];
                    if (aij == 0) continue;
                    double th = (A[j * 3 + j] - A[i * 3 + i]) / (2 * aij);
                    double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1));
                    double cc = 1 / sqrt(t * t + 1), ss = t * cc;
#pragma unroll
                    for (int k = 0; k < p; ++k) {
                        double aki = A[k * 3 + i], akj = A[k * 3 + j];
                        A[k * 3 + i] = cc * aki - ss * akj;
                        A[k * 3 + j] = ss * aki + cc * akj;
                    }
#pragma unroll
                    for (int k = 0; k < p; ++k) {
                        double aik = A[i * 3 + k], ajk = A[j * 3 + k];
                        A[i * 3 + k] = cc * aik - ss * ajk;
                        A[j * 3 + k] = ss * aik + cc * ajk;
                    }
#pragma unroll
                    for (int k = 0; k < p; ++k) {
                        double vki = E[k * 3 + i], vkj = E[k * 3 + j];
                        E[k * 3 + i] = cc * vki - ss * vkj;
                        E[k * 3 + j] = ss * vki + cc * vkj;
                    }
                }
        }
        double smax = 0;
#pragma unroll
        for (int i = 0; i < p; ++i) smax = fmax(smax, fabs(A[i * 3 + i]));
        double tol = DEPS * p * smax;
#pragma unroll
        for (int i = 0; i < p; ++i)
#pragma unroll
            for (int j = 0; j < p; ++j) {
                double s = 0;
#pragma unroll
                for (int k = 0; k < p; ++k)
                    if (fabs(A[k * 3 + k]) > tol) s += E[i * 3 + k] * E[j * 3 + k] / A[k * 3 + k];
                cov[i * 3 + j] = s;
            }
    }
    double fac = cov_mode == 2 ? fac_rw : 1.0;
    if (sub == 0)
#pragma unroll
        for (int k = 0; k < p; ++k) errs[(size_t)a * p + k] = sqrt(fmax(0.0, fac * cov[k * 3 + k]));
}

// Post-fit per-star statistics over the +-fit_rad box at the fitted position
// (psf_photometry_simultaneous.jl:1344-1397; _star_diagnostics!,
// psf_photometry_diagnostics.jl:154-231): one wave per star.  The star's
// model over the box is staged in LDS and correlated (zero border) with the
// spread_model reference disk K of half-width kh (K = null: spread_model
// stays NaN).  resid is the global residual (data minus every star's model),
// data the working image, w the weights (inverse variance, 0 = masked; all
// ones without inv_var).  Valid stars write the result columns OUT_CHISQ ..
// OUT_CHISQ + 6 (chisq, qfit, qfit_expected, qfit_z, crowding, spread_model,
// spread_model_err; NaN where the reference leaves NaN) of catalogue star
// act[a]; invalid stars keep what k_results wrote (chisq 0, the rest NaN).
constexpr int DIAG_GL = 16;      // lanes per star in k_diag
constexpr int DIAG_MAXB = 1024;   // box pixels: (2 ceil(fit_rad) + 2)^2 for fit_rad <= 15

// Result columns (float64, catalogue order, n each), then n_iter (int32),
// converged (u8) and valid (u8): the per-star fields of MultiPassPhotResult
// (src/photometry/psf_photometry_single.jl:99-122).
enum OutCol {
    OUT_Y, OUT_X, OUT_FLUX, OUT_BKG, OUT_Y_ERR, OUT_X_ERR, OUT_FLUX_ERR, OUT_BKG_ERR, OUT_CHISQ,
    OUT_QFIT, OUT_QFIT_EXP, OUT_QFIT_Z, OUT_CROWDING, OUT_SPREAD, OUT_SPREAD_ERR, OUT_F64
};

// Live compaction of the model-box tile lists: flag[i] = live[list[i]] for
// i < n, flag[n] = 0.
static __global__ void k_list_live(int n, const int* __restrict__ list, const uint8_t* __restrict__ live,
                                   int* __restrict__ flag) {
    for (int i = blockIdx.x * BLOCK + threadIdx.x; i <= n; i += gridDim.x * BLOCK)
        flag[i] = i < n ? (int)live[list[i]] : 0;
}

// list_l[pos[i]] = list[i] for flagged entries (pos = exclusive scan of flag)
// and the tiles' ranges in the compacted list.
static __global__ void k_list_compact(int n, const int* __restrict__ list, const int* __restrict__ flag,
                                      const int* __restrict__ pos, int* __restrict__ list_l, int ntiles,
                                      const int* __restrict__ start, const int* __restrict__ end,
                                      int* __restrict__ start_l, int* __restrict__ end_l) {
    const int stride = gridDim.x * BLOCK;
    for (int i = blockIdx.x * BLOCK + threadIdx.x; i < n; i += stride)
        if (flag[i]) list_l[pos[i]] = list[i];
    for (int t = blockIdx.x * BLOCK + threadIdx.x; t < ntiles; t += stride) {
        start_l[t] = pos[start[t]];
        end_l[t] = pos[end[t]];
    }
}

// The single-node ePSF grid in the edge-padded polyphase layout the LDS
// copies use (built once per fit; kernels copy it linearly).
static __global__ void k_poly_grid(const float* __restrict__ g, int gny, int gnx, int osy, int osx, int Qy, int Qx,
                                   float* __restrict__ out) {
    stage_poly(g, gny, gnx, osy, osx, Qy, Qx, out, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

// slot[act[j]] = j: the active index of each catalogue star (slot is -1
// beforehand for every star).
static __global__ void k_slot_of(const int* __restrict__ act, int na, int* __restrict__ slot) {
    const int j = blockIdx.x * BLOCK + threadIdx.x;
    if (j < na) slot[act[j]] = j;
}

// Per-catalogue-star results (:1289-1300, :1331-1337, :1414-1430): an active
// star (slot >= 0) takes its fitted free parameters and, with err_on, their
// errors; every other value is the catalogue's (errors NaN for free, 0 for
// fixed parameters); fixed parameters take their fixed value; valid = active
// and finite y, x and positive finite flux (the final validity gate); n_iter
// = n_run for a star still live at the end, else its freeze iteration.
// chisq starts at 0 and the diagnostics at NaN (k_diag fills valid stars).
static __global__ void k_results(int n, const int* __restrict__ slot, const double* __restrict__ y0,
                                 const double* __restrict__ x0, const double* __restrict__ f0,
                                 const double* __restrict__ theta, const double* __restrict__ errs, int err_on,
                                 const uint8_t* __restrict__ live, const uint8_t* __restrict__ conv,
                                 const int* __restrict__ frozen_at, int n_run, FreeMap fm, double fixed_bkg,
                                 double* __restrict__ cols, int* __restrict__ n_iter, uint8_t* __restrict__ converged,
                                 uint8_t* __restrict__ valid) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const int a = slot[i];
    double Y = y0[i], X = x0[i], F = f0[i];
    double ey = fm.y_fixed ? 0.0 : NAN, ex = fm.x_fixed ? 0.0 : NAN, ef = fm.f_fixed ? 0.0 : NAN;
    int it = 0;
    uint8_t cv = 0;
    if (a >= 0) {
        for (int k = 0; k < fm.p; ++k) {
            const int c = gc_of(fm, k);
            const double v = theta[(size_t)a * fm.p + k], e = errs[(size_t)a * fm.p + k];
            if (c == 1) { Y = v; if (err_on) ey = e; }
            if (c == 2) { X = v; if (err_on) ex = e; }
            if (c == 3) { F = v; if (err_on) ef = e; }
        }
        cv = conv[a];
        it = live[a] ? n_run : frozen_at[a];
    }
    if (fm.y_fixed) Y = fm.fixed_y;
    if (fm.x_fixed) X = fm.fixed_x;
    if (fm.f_fixed) F = fm.fixed_f;
    const size_t N = (size_t)n;
    cols[OUT_Y * N + i] = Y;
    cols[OUT_X * N + i] = X;
    cols[OUT_FLUX * N + i] = F;
    cols[OUT_BKG * N + i] = fixed_bkg;
    cols[OUT_Y_ERR * N + i] = ey;
    cols[OUT_X_ERR * N + i] = ex;
    cols[OUT_FLUX_ERR * N + i] = ef;
    cols[OUT_BKG_ERR * N + i] = 0.0;
    cols[OUT_CHISQ * N + i] = 0.0;
    for (int k = OUT_QFIT; k < OUT_F64; ++k) cols[k * N + i] = NAN;
    n_iter[i] = it;
    converged[i] = cv;
    valid[i] = a >= 0 && isfinite(Y) && isfinite(X) && isfinite(F) && F > 0;
}

template <typename T, int KIND>
static __global__ __launch_bounds__(64) void k_diag(int na, const double* __restrict__ theta, FreeMap fm,
                                                    double fit_rad, ModelConst<T> M, double inv_norm_g64,
                                                    double moffat_amp64, NodeGrid G, const T* __restrict__ resid,
                                                    const T* __restrict__ data, const T* __restrict__ w,
                                                    int has_iv, int ny, int nx, const double* __restrict__ K,
                                                    int kh, const int* __restrict__ act, int n, int maxb,
                                                    double* __restrict__ cols) {
    // DIAG_GL lanes per star, 64 / DIAG_GL stars per block, each with its
    // own box of model values in LDS (maxb entries).  The only barrier comes
    // before any lane leaves; group sums stay inside a group.
    extern __shared__ __align__(16) unsigned char dsm[];
    const int grp = threadIdx.x / DIAG_GL, lane = threadIdx.x % DIAG_GL;
    const int a = blockIdx.x * (64 / DIAG_GL) + grp;
    T* sm = reinterpret_cast<T*>(dsm) + grp * maxb;
    const size_t N = (size_t)n;
    double Y = 0, X = 0, F = 0;
    if (a < na) star_yxf(fm, theta + (size_t)a * fm.p, Y, X, F);
    const bool valid = a < na && isfinite(Y) && isfinite(X) && isfinite(F) && F > 0;
    int ylo = 1, yhi = 0, xlo = 1, xhi = 0;
    if (valid) {
        ylo = max((int)floor(Y - fit_rad), 1);
        yhi = min((int)ceil(Y + fit_rad), ny);
        xlo = max((int)floor(X - fit_rad), 1);
        xhi = min((int)ceil(X + fit_rad), nx);
    }
    const int by = yhi - ylo + 1, bx = xhi - xlo + 1;
    const bool go = valid && by > 0 && bx > 0;
    const int nb = go ? by * bx : 0;
    StarRec<T> r{};
    r.flux = (T)F;
    if (KIND == KIND_GAUSS) r.amp = (T)(F * inv_norm_g64);
    else if (KIND == KIND_MOFFAT) r.amp = (T)(F * moffat_amp64);
    else r.amp = (T)F;
    CornerRec<T> cr;
    if (KIND == KIND_EPSF && go) grid_corners<T>(G, M.n_nodes, Y, X, cr);
    for (int m = lane; m < nb; m += DIAG_GL) {
        const int iy = m % by, ix = m / by;
        sm[m] = star_value<T, KIND>(M, r, (T)((ylo + iy) - Y), (T)((xlo + ix) - X), M.grids, &cr);
    }
    __syncthreads();
    if (!go) return;
    double* o = cols + OUT_CHISQ * N + act[a];   // o[k * N]: column OUT_CHISQ + k
    const double bkg = (double)M.bkg, inv_flux = 1.0 / F;
    double chi = 0, qfit = 0, nclean = 0, ndirty = 0, dcrowd = 0, Sp = 0, Sq = 0, Sp0 = 0, Sq0 = 0, SqG = 0;
    double ssum = 0, s2sum = 0, ngood = 0;
    const int kw = 2 * kh + 1;
    for (int m = lane; m < nb; m += DIAG_GL) {
        const int iy = m % by, ix = m / by;
        const size_t q = (size_t)(ylo + iy - 1) + (size_t)(xlo + ix - 1) * ny;
        const double wp = (double)w[q];
        if (!(wp > 0)) continue;
        const double rq = (double)resid[q], mv = (double)sm[m];
        chi += wp * rq * rq;
        ngood += 1;
        qfit += fabs(rq);
        const double Pp = (mv - bkg) * inv_flux, wP = wp * Pp, pi = rq + mv - bkg;
        nclean += wP * pi;
        ndirty += wP * ((double)data[q] - bkg);
        dcrowd += wP * Pp;
        if (K) {
            double gm = 0;
            for (int bb = -kh; bb <= kh; ++bb) {
                const int jx = ix + bb;
                if (jx < 0 || jx >= bx) continue;
                for (int aa = -kh; aa <= kh; ++aa) {
                    const int jy = iy + aa;
                    if (jy < 0 || jy >= by) continue;
                    gm += (double)sm[jy + jx * by] * K[(aa + kh) + (bb + kh) * kw];
                }
            }
            const double gi = (gm - bkg) * inv_flux;
            Sp += wP * pi;
            Sq += wp * gi * pi;
            Sp0 += wP * Pp;
            Sq0 += wp * Pp * gi;
            SqG += wp * gi * gi;
        }
        if (has_iv) {
            const double sg = 1.0 / sqrt(wp);
            ssum += sg;
            s2sum += sg * sg;
        }
    }
    chi = group_sum<DIAG_GL>(chi);
    ngood = group_sum<DIAG_GL>(ngood);
    qfit = group_sum<DIAG_GL>(qfit);
    nclean = group_sum<DIAG_GL>(nclean);
    ndirty = group_sum<DIAG_GL>(ndirty);
    dcrowd = group_sum<DIAG_GL>(dcrowd);
    Sp = group_sum<DIAG_GL>(Sp);
    Sq = group_sum<DIAG_GL>(Sq);
    Sp0 = group_sum<DIAG_GL>(Sp0);
    Sq0 = group_sum<DIAG_GL>(Sq0);
    SqG = group_sum<DIAG_GL>(SqG);
    ssum = group_sum<DIAG_GL>(ssum);
    s2sum = group_sum<DIAG_GL>(s2sum);
    if (lane != 0) return;
    const int p = fm.p;
    o[0] = ngood - p > 0 ? chi / (ngood - p) : 0.0;
    o[N] = qfit * inv_flux;
    o[4 * N] = (dcrowd > 0 && nclean > 0 && ndirty > 0) ? 2.5 * log10(ndirty / nclean) : NAN;
    o[5 * N] = (K && Sp > 0) ? Sq / Sp - Sq0 / Sp0 : NAN;
    if (K && Sp > 0 && has_iv) {
        const double v = Sp * Sp * SqG + Sq * Sq * Sp0 - 2 * Sq * Sp * Sq0;
        o[6 * N] = sqrt(fmax(0.0, v)) / (Sp * Sp);
    } else {
        o[6 * N] = NAN;
    }
    const double c2pi = sqrt(2.0 / M_PI);
    o[2 * N] = has_iv ? c2pi * ssum * inv_flux : NAN;
    if (has_iv && s2sum > 0 && ngood > p) {
        const double dof = sqrt(1.0 - p / ngood);
        o[3 * N] = (qfit - c2pi * dof * ssum) / (sqrt((1 - 2 / M_PI) * s2sum) * dof);
    } else {
        o[3 * N] = NAN;
    }
}

// dst (cols x rows, row-major) = transpose of src (rows x cols, row-major):
// turns a C-order image into the column-major layout and back.  32 x 32 LDS
// tiles (one padding column against bank conflicts), block 32 x 8.
template <typename TS, typename TD>
static __global__ __launch_bounds__(256) void k_transpose(const TS* __restrict__ src, int rows, int cols,
                                                          TD* __restrict__ dst) {
    __shared__ TD tile[32][33];
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    for (int i = threadIdx.y; i < 32; i += 8) {
        const int r = r0 + i, c = c0 + threadIdx.x;
        if (r < rows && c < cols) tile[i][threadIdx.x] = (TD)src[(size_t)r * cols + c];
    }
    __syncthreads();
    for (int i = threadIdx.y; i < 32; i += 8) {
        const int c = c0 + i, r = r0 + threadIdx.x;
        if (r < rows && c < cols) dst[(size_t)c * rows + r] = tile[threadIdx.x][i];
    }
}

template <typename T>
static __global__ void k_to_float(const T* __restrict__ src, float* __restrict__ dst, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK)
        dst[i] = (float)src[i];
}

static __global__ void k_w_sample(const float* __restrict__ src, int64_t P, int64_t step, float* __restrict__ dst,
                           int64_t nout) {
    int64_t i = blockIdx.x * (int64_t)BLOCK + threadIdx.x;
    if (i < nout) dst[i] = src[i * step];
}

}  // namespace cp
