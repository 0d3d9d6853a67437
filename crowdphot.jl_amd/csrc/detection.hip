// detection.hip -- correlate, findlocalmaxima and matched_filter on gfx950
// (src/correlation.jl:108-547, src/detection.jl:87-324): the catalogue step
// in front of the fits.
//
// Correlation.  A kernel the reference would factor (numerically rank 1,
// _tryfactor correlation.jl:196-226) runs as a column pass then a row pass,
// each thread one output pixel reading its taps straight from L1/L2 (the
// reads of one tap are coalesced down a column).  Any other kernel runs as a
// 2-D pass over (16 RY) x 16 output tiles: the block stages the tile's input
// window (border rule applied) in LDS as float64, the kernel taps come
// through the scalar cache, and every thread sums RY output rows, one LDS
// window read feeding RY FMAs; windows that would not fit LDS read the image
// through the cache instead.  Sums are float64 in kernel order.
//
// Peaks.  One thread per pixel tests the strict 3 x 3 maximum; the flagged
// pixels are compacted in the reference's order (column-major interior,
// then the edge groups) with hipcub.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

#include "cp_internal.h"
#include "psf_host.h"

namespace cp {
namespace {

constexpr int CT = 16;                   // output tile side of the 2-D pass
constexpr size_t CORR_LDS_MAX = 64 * 1024;

template <typename TI>
__device__ __forceinline__ double pix(const TI* __restrict__ img, int ny, int nx, int r, int c, int border) {
    if (border == CP_BORDER_REPLICATE) {
        r = min(max(r, 0), ny - 1);
        c = min(max(c, 0), nx - 1);
        return (double)img[(size_t)r + (size_t)c * ny];
    }
    return (r >= 0 && r < ny && c >= 0 && c < nx) ? (double)img[(size_t)r + (size_t)c * ny] : 0.0;
}

// 2-D pass, window staged in LDS.  Block CT x CT threads (y fastest); each
// thread sums RY consecutive output rows of one column, so one LDS read of
// the window feeds RY fused multiply-adds: for each kernel column the RY
// running window values rotate through registers (the i loop is unrolled by
// RY so the rotation is register renaming).  Kernel taps are wave-uniform
// loads through the scalar cache.  Each output sums its taps column by
// column, rows ascending.
template <typename TI, int RY>
__global__ __launch_bounds__(CT* CT) void k_corr2d_lds(const TI* __restrict__ img, int ny, int nx,
                                                       const double* __restrict__ K, int kr, int kc, int border,
                                                       double* __restrict__ out) {
    extern __shared__ double win[];
    constexpr int TY = CT * RY;
    const int r = kr / 2, c = kc / 2;
    const int wy = TY + kr - 1, wx = CT + kc - 1;
    const int row0 = blockIdx.x * TY, col0 = blockIdx.y * CT;
    for (int i = threadIdx.x; i < wy * wx; i += CT * CT) {
        const int b = i / wy, a = i - b * wy;
        win[i] = pix(img, ny, nx, row0 + a - r, col0 + b - c, border);
    }
    __syncthreads();
    const int ty = threadIdx.x % CT, tx = threadIdx.x / CT;
    const int rb = ty * RY, col = col0 + tx;
    double acc[RY];
#pragma unroll
    for (int q = 0; q < RY; ++q) acc[q] = 0.0;
    const int kmain = kr - kr % RY;
    for (int j = 0; j < kc; ++j) {
        const double* w = win + rb + (size_t)(tx + j) * wy;
        const double* kj = K + (size_t)j * kr;
        double v[RY];
#pragma unroll
        for (int q = 0; q < RY - 1; ++q) v[q] = w[q];
        int i = 0;
        for (; i < kmain; i += RY) {
#pragma unroll
            for (int u = 0; u < RY; ++u) {
                v[(u + RY - 1) % RY] = w[i + u + RY - 1];
                const double k = kj[i + u];
#pragma unroll
                for (int q = 0; q < RY; ++q) acc[q] = fma(v[(u + q) % RY], k, acc[q]);
            }
        }
        // remainder taps (kr not a multiple of RY): rebuild the window per tap
        for (; i < kr; ++i) {
            const double k = kj[i];
#pragma unroll
            for (int q = 0; q < RY; ++q) acc[q] = fma(w[i + q], k, acc[q]);
        }
    }
    if (col >= nx) return;
#pragma unroll
    for (int q = 0; q < RY; ++q) {
        const int row = row0 + rb + q;
        if (row < ny) out[(size_t)row + (size_t)col * ny] = acc[q];
    }
}

template <int RY>
constexpr size_t corr_lds_bytes(int kr, int kc) {
    return (size_t)(CT * RY + kr - 1) * (CT + kc - 1) * 8;
}

// matched_filter's three correlations in one pass: smoothed =
// correlate(image, K) and, with weights, num = correlate(w v, K) and
// den = correlate(w, K^2), then the significance num / sqrt(den) (0 where
// den <= 0) or smoothed * knorm.  The block stages the image and (weighted)
// the weights w = finite(inv_var) ? inv_var : 0 as float32 planes (exact
// copies of the float32 inputs, border rule applied), and each window value
// read feeds RY outputs of every sum; w v is formed in float64 at read time
// (exact for float32 operands; 0 where w = 0).  Same tap order as correlate.
template <int RY, bool WEIGHTED>
__global__ __launch_bounds__(CT* CT) void k_mf2d(const float* __restrict__ img, const float* __restrict__ iv, int ny,
                                                 int nx, const double* __restrict__ K, const double* __restrict__ K2,
                                                 int kr, int kc, int border, double knorm, double* __restrict__ sm,
                                                 double* __restrict__ den_out, double* __restrict__ sig) {
    extern __shared__ float fsh[];
    constexpr int TY = CT * RY;
    const int r = kr / 2, c = kc / 2;
    const int wy = TY + kr - 1, wx = CT + kc - 1;
    float* pv = fsh;
    float* pw = fsh + (size_t)wy * wx;
    const int row0 = blockIdx.x * TY, col0 = blockIdx.y * CT;
    for (int i = threadIdx.x; i < wy * wx; i += CT * CT) {
        const int b = i / wy, a = i - b * wy;
        int rr = row0 + a - r, cc = col0 + b - c;
        bool in = true;
        if (border == CP_BORDER_REPLICATE) {
            rr = min(max(rr, 0), ny - 1);
            cc = min(max(cc, 0), nx - 1);
        } else {
            in = rr >= 0 && rr < ny && cc >= 0 && cc < nx;
        }
        const size_t q = in ? (size_t)rr + (size_t)cc * ny : 0;
        pv[i] = in ? img[q] : 0.0f;
        if (WEIGHTED) {
            const float a_ = in ? iv[q] : 0.0f;
            pw[i] = isfinite(a_) ? a_ : 0.0f;
        }
    }
    __syncthreads();
    const int ty = threadIdx.x % CT, tx = threadIdx.x / CT;
    const int rb = ty * RY, col = col0 + tx;
    double as[RY], an[RY], ad[RY];
#pragma unroll
    for (int q = 0; q < RY; ++q) as[q] = an[q] = ad[q] = 0.0;
    // one tap for all RY outputs from the rotating registers xv / xw / xp
    // (image, weight, weighted image) at rotation u
    auto tap = [&](const double* xv, const double* xw, const double* xp, int u, double k, double k2) {
#pragma unroll
        for (int q = 0; q < RY; ++q) {
            const int e = (u + q) % RY;
            as[q] = fma(xv[e], k, as[q]);
            if (WEIGHTED) {
                an[q] = fma(xp[e], k, an[q]);
                ad[q] = fma(xw[e], k2, ad[q]);
            }
        }
    };
    const int kmain = kr - kr % RY;
    for (int j = 0; j < kc; ++j) {
        const float* v = pv + rb + (size_t)(tx + j) * wy;
        const float* w = pw + rb + (size_t)(tx + j) * wy;
        const double* kj = K + (size_t)j * kr;
        const double* k2j = K2 + (size_t)j * kr;
        double xv[RY], xw[RY], xp[RY];
        auto load = [&](int slot, int at) {
            xv[slot] = (double)v[at];
            if (WEIGHTED) {
                xw[slot] = (double)w[at];
                xp[slot] = xw[slot] == 0.0 ? 0.0 : xw[slot] * xv[slot];
            }
        };
#pragma unroll
        for (int q = 0; q < RY - 1; ++q) load(q, q);
        int i = 0;
        for (; i < kmain; i += RY) {
#pragma unroll
            for (int u = 0; u < RY; ++u) {
                load((u + RY - 1) % RY, i + u + RY - 1);
                tap(xv, xw, xp, u, kj[i + u], WEIGHTED ? k2j[i + u] : 0.0);
            }
        }
        for (; i < kr; ++i) {   // remainder taps: reload the RY window values
#pragma unroll
            for (int q = 0; q < RY; ++q) load(q, i + q);
            tap(xv, xw, xp, 0, kj[i], WEIGHTED ? k2j[i] : 0.0);
        }
    }
    if (col >= nx) return;
#pragma unroll
    for (int q = 0; q < RY; ++q) {
        const int row = row0 + rb + q;
        if (row >= ny) continue;
        const size_t o = (size_t)row + (size_t)col * ny;
        sm[o] = as[q];
        if (WEIGHTED) {
            den_out[o] = ad[q];
            sig[o] = ad[q] > 0 ? an[q] / sqrt(ad[q]) : 0.0;
        } else {
            sig[o] = as[q] * knorm;
        }
    }
}

template <int RY>
constexpr size_t mf_lds_bytes(int kr, int kc, bool weighted) {
    return (size_t)(CT * RY + kr - 1) * (CT + kc - 1) * 4 * (weighted ? 2 : 1);
}

// 2-D pass through the cache, for kernels whose window overflows LDS.
template <typename TI>
__global__ __launch_bounds__(BLOCK) void k_corr2d_global(const TI* __restrict__ img, int ny, int nx,
                                                         const double* __restrict__ K, int kr, int kc, int border,
                                                         double* __restrict__ out) {
    const int64_t q = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (q >= (int64_t)ny * nx) return;
    const int col = (int)(q / ny), row = (int)(q - (int64_t)col * ny);
    const int r = kr / 2, c = kc / 2;
    double acc = 0.0;
    for (int j = 0; j < kc; ++j)
        for (int i = 0; i < kr; ++i)
            acc = fma(pix(img, ny, nx, row + i - r, col + j - c, border), K[i + (size_t)j * kr], acc);
    out[q] = acc;
}

// 1-D pass along rows (dimension 1, an m x 1 factor) or columns (1 x n).
template <typename TI, bool ALONG_ROWS>
__global__ __launch_bounds__(BLOCK) void k_corr1d(const TI* __restrict__ img, int ny, int nx,
                                                  const double* __restrict__ k, int len, int border,
                                                  double* __restrict__ out) {
    const int64_t q = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (q >= (int64_t)ny * nx) return;
    const int col = (int)(q / ny), row = (int)(q - (int64_t)col * ny);
    const int h = len / 2;
    double acc = 0.0;
    for (int t = 0; t < len; ++t)
        acc = ALONG_ROWS ? fma(pix(img, ny, nx, row + t - h, col, border), k[t], acc)
                         : fma(pix(img, ny, nx, row, col + t - h, border), k[t], acc);
    out[q] = acc;
}

// Order key of a strict 3 x 3 local maximum in the reference's scan order
// (interior column-major, then top / bottom rows by column, left / right
// columns by row, then the corners), or ~0 when the pixel is not one.
// interior_only: edges = false.  min_value: also require v >= min_value.
__device__ __forceinline__ unsigned long long max_key(const double* __restrict__ a, int H, int W, int row, int col,
                                                      bool interior_only, double min_value) {
    auto v = [&](int r, int c) { return a[(size_t)(r - 1) + (size_t)(c - 1) * H]; };
    const double x = v(row, col);
    const unsigned long long none = ~0ull;
    if (!(x >= min_value)) return none;
    const unsigned long long P = (unsigned long long)H * W, Wl = W, Hl = H;
    const bool in_r = row >= 2 && row <= H - 1, in_c = col >= 2 && col <= W - 1;
    if (in_r && in_c) {
        const bool m = x > v(row - 1, col - 1) && x > v(row - 1, col) && x > v(row - 1, col + 1) &&
                       x > v(row, col - 1) && x > v(row, col + 1) && x > v(row + 1, col - 1) &&
                       x > v(row + 1, col) && x > v(row + 1, col + 1);
        return m ? (unsigned long long)(row - 1) + (unsigned long long)(col - 1) * Hl : none;
    }
    if (interior_only) return none;
    if (in_c && (row == 1 || row == H)) {
        const int n = row == 1 ? 2 : H - 1;
        const bool m = x > v(row, col - 1) && x > v(row, col + 1) && x > v(n, col - 1) && x > v(n, col) &&
                       x > v(n, col + 1);
        return m ? P + 2ull * (col - 2) + (row == 1 ? 0 : 1) : none;
    }
    if (in_r && (col == 1 || col == W)) {
        const int n = col == 1 ? 2 : W - 1;
        const bool m = x > v(row - 1, col) && x > v(row - 1, n) && x > v(row, n) && x > v(row + 1, col) &&
                       x > v(row + 1, n);
        return m ? P + 2ull * Wl + 2ull * (row - 2) + (col == 1 ? 0 : 1) : none;
    }
    if (H >= 2 && W >= 2 && (row == 1 || row == H) && (col == 1 || col == W)) {
        const int nr = row == 1 ? 2 : H - 1, nc = col == 1 ? 2 : W - 1;
        const bool m = x > v(row, nc) && x > v(nr, col) && x > v(nr, nc);
        const int corner = (row == 1 ? 0 : 2) + (col == 1 ? 0 : 1);
        return m ? P + 2ull * Wl + 2ull * Hl + corner : none;
    }
    return none;
}

static __global__ void k_max_keys(const double* __restrict__ a, int H, int W, int interior_only, double min_value,
                                  unsigned long long* __restrict__ key, uint8_t* __restrict__ flag) {
    const int64_t q = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (q >= (int64_t)H * W) return;
    const int col = (int)(q / H), row = (int)(q - (int64_t)col * H);
    const unsigned long long k = max_key(a, H, W, row + 1, col + 1, interior_only != 0, min_value);
    key[q] = k;
    flag[q] = k != ~0ull;
}

// Peak rows / columns (1-based) and, for matched_filter, significance and flux.
static __global__ void k_peak_out(int n, const int* __restrict__ q, int H, const double* __restrict__ sig,
                                  const double* __restrict__ sm, const double* __restrict__ den, double kn2,
                                  int* __restrict__ py, int* __restrict__ px, double* __restrict__ ps,
                                  double* __restrict__ pf) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const int p = q[i];
    py[i] = p % H + 1;
    px[i] = p / H + 1;
    if (!sig) return;
    const double s = sig[p];
    ps[i] = s;
    pf[i] = den ? (den[p] > 0 ? s / (sqrt(den[p]) * kn2) : 0.0) : sm[p];
}

// matched_filter's weight products: w = finite(inv_var) ? inv_var : 0 and
// w v (0 where w = 0, so a NaN under a masked pixel does not leak).
static __global__ void k_mf_weights(int64_t P, const float* __restrict__ v, const float* __restrict__ iv,
                                    double* __restrict__ w, double* __restrict__ wv) {
    const int64_t q = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (q >= P) return;
    const double a = (double)iv[q];
    const double ww = isfinite(a) ? a : 0.0;
    w[q] = ww;
    wv[q] = ww == 0.0 ? 0.0 : ww * (double)v[q];
}

static __global__ void k_mf_significance(int64_t P, const double* __restrict__ num, const double* __restrict__ den,
                                         const double* __restrict__ sm, double knorm, double* __restrict__ sig) {
    const int64_t q = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (q >= P) return;
    sig[q] = den ? (den[q] > 0 ? num[q] / sqrt(den[q]) : 0.0) : sm[q] * knorm;
}

// One-sided Jacobi SVD of the kr x kc kernel: true when every singular value
// past the first is below sqrt(eps), with the factors sqrt(s1) u1 (kr) and
// sqrt(s1) v1 (kc) (_tryfactor, correlation.jl:196-226).
bool factor_rank1(const double* K, int kr, int kc, std::vector<double>& k1, std::vector<double>& k2) {
    if (kr == 1 && kc == 1) return false;
    // A (m x n, m >= n) = K or K'; columns orthogonalised in place, V accumulated
    const bool tr = kc > kr;
    const int m = tr ? kc : kr, n = tr ? kr : kc;
    std::vector<double> A((size_t)m * n), V((size_t)n * n, 0.0);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) A[i + (size_t)j * m] = tr ? K[j + (size_t)i * kr] : K[i + (size_t)j * kr];
    for (int j = 0; j < n; ++j) V[j + (size_t)j * n] = 1.0;
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                double a = 0, b = 0, g = 0;
                for (int i = 0; i < m; ++i) {
                    const double x = A[i + (size_t)p * m], y = A[i + (size_t)q * m];
                    a += x * x;
                    b += y * y;
                    g += x * y;
                }
                if (g == 0 || std::fabs(g) <= 1e-300) continue;
                off = std::max(off, std::fabs(g) / std::sqrt(a * b));
                const double zeta = (b - a) / (2 * g);
                const double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1 + zeta * zeta));
                const double cs = 1 / std::sqrt(1 + t * t), sn = cs * t;
                for (int i = 0; i < m; ++i) {
                    const double x = A[i + (size_t)p * m], y = A[i + (size_t)q * m];
                    A[i + (size_t)p * m] = cs * x - sn * y;
                    A[i + (size_t)q * m] = sn * x + cs * y;
                }
                for (int i = 0; i < n; ++i) {
                    const double x = V[i + (size_t)p * n], y = V[i + (size_t)q * n];
                    V[i + (size_t)p * n] = cs * x - sn * y;
                    V[i + (size_t)q * n] = sn * x + cs * y;
                }
            }
        if (off < 1e-15) break;
    }
    std::vector<double> s(n);
    int jmax = 0;
    for (int j = 0; j < n; ++j) {
        double a = 0;
        for (int i = 0; i < m; ++i) a += A[i + (size_t)j * m] * A[i + (size_t)j * m];
        s[j] = std::sqrt(a);
        if (s[j] > s[jmax]) jmax = j;
    }
    const double EPS = std::sqrt(2.220446049250313e-16);
    for (int j = 0; j < n; ++j)
        if (j != jmax && !(s[j] < EPS)) return false;
    const double s1 = s[jmax], ss = std::sqrt(s1);
    if (!(s1 > 0)) return false;
    std::vector<double> u(m), v(n);
    for (int i = 0; i < m; ++i) u[i] = A[i + (size_t)jmax * m] / s1;
    for (int i = 0; i < n; ++i) v[i] = V[i + (size_t)jmax * n];
    // K = s1 u v' (or its transpose)
    k1.resize(kr);
    k2.resize(kc);
    for (int i = 0; i < kr; ++i) k1[i] = ss * (tr ? v[i] : u[i]);
    for (int j = 0; j < kc; ++j) k2[j] = ss * (tr ? u[j] : v[j]);
    return true;
}

struct CorrPlan {
    bool sep = false;
    DevBuf K, k1, k2;
    int kr = 0, kc = 0;
};

void plan_corr(cp_ctx* ctx, const double* K, int kr, int kc, CorrPlan& pl) {
    CP_ARG(kr > 0 && kc > 0 && kr % 2 == 1 && kc % 2 == 1,
           "kernel dimensions must be odd (got " + std::to_string(kr) + "x" + std::to_string(kc) +
               "); the correlation center is at ((m+1)/2, (n+1)/2), which requires odd-sized kernels");
    hipStream_t st = ctx->stream;
    pl.kr = kr;
    pl.kc = kc;
    std::vector<double> a, b;
    pl.sep = factor_rank1(K, kr, kc, a, b);
    if (pl.sep) {
        pl.k1.alloc(ctx, (size_t)kr * 8);
        pl.k2.alloc(ctx, (size_t)kc * 8);
        CP_HIP(hipMemcpyAsync(pl.k1.get(), a.data(), (size_t)kr * 8, hipMemcpyHostToDevice, st));
        CP_HIP(hipMemcpyAsync(pl.k2.get(), b.data(), (size_t)kc * 8, hipMemcpyHostToDevice, st));
        CP_HIP(hipStreamSynchronize(st));   // a, b are pageable and local
    } else {
        pl.K.alloc(ctx, (size_t)kr * kc * 8);
        CP_HIP(hipMemcpyAsync(pl.K.get(), K, (size_t)kr * kc * 8, hipMemcpyHostToDevice, st));
    }
}

template <typename TI>
void run_corr(cp_ctx* ctx, const CorrPlan& pl, const TI* d_img, int ny, int nx, int border, double* d_out,
              DevBuf& tmp) {
    hipStream_t st = ctx->stream;
    const int64_t P = (int64_t)ny * nx;
    const unsigned g = (unsigned)((P + BLOCK - 1) / BLOCK);
    if (pl.sep) {
        if (tmp.bytes() < (size_t)P * 8) tmp.alloc(ctx, (size_t)P * 8);
        launch_n("k_corr1d", k_corr1d<TI, true>, dim3(g), dim3(BLOCK), 0, st, d_img, ny, nx,
                 (const double*)pl.k1.as<double>(), pl.kr, border, tmp.as<double>());
        launch_n("k_corr1d", k_corr1d<double, false>, dim3(g), dim3(BLOCK), 0, st, (const double*)tmp.as<double>(),
                 ny, nx, (const double*)pl.k2.as<double>(), pl.kc, border, d_out);
        return;
    }
    auto go = [&](auto ry) {
        constexpr int RY = decltype(ry)::value;
        dim3 gg((unsigned)((ny + CT * RY - 1) / (CT * RY)), (unsigned)((nx + CT - 1) / CT));
        launch_n("k_corr2d", k_corr2d_lds<TI, RY>, gg, dim3(CT * CT), corr_lds_bytes<RY>(pl.kr, pl.kc), st, d_img,
                 ny, nx, (const double*)pl.K.as<double>(), pl.kr, pl.kc, border, d_out);
    };
    if (corr_lds_bytes<8>(pl.kr, pl.kc) <= CORR_LDS_MAX) {
        go(std::integral_constant<int, 8>{});
    } else if (corr_lds_bytes<2>(pl.kr, pl.kc) <= CORR_LDS_MAX) {
        go(std::integral_constant<int, 2>{});
    } else {
        launch_n("k_corr2d", k_corr2d_global<TI>, dim3(g), dim3(BLOCK), 0, st, d_img, ny, nx,
                 (const double*)pl.K.as<double>(), pl.kr, pl.kc, border, d_out);
    }
}

// Keys of the maxima of d_a (n_pix), compacted and sorted: *count and the
// 0-based linear indices in the reference's order in d_idx.
int select_maxima(cp_ctx* ctx, const double* d_a, int H, int W, bool interior_only, double min_value,
                  DevBuf& d_idx) {
    hipStream_t st = ctx->stream;
    const int64_t P = (int64_t)H * W;
    const unsigned g = (unsigned)((P + BLOCK - 1) / BLOCK);
    DevBuf key(ctx, (size_t)P * 8), flag(ctx, (size_t)P), sel_k(ctx, (size_t)P * 8), sel_q(ctx, (size_t)P * 4),
        nsel(ctx, 8), tmp;
    launch_n("k_max_keys", k_max_keys, dim3(g), dim3(BLOCK), 0, st, d_a, H, W, interior_only ? 1 : 0, min_value,
             key.as<unsigned long long>(), flag.as<uint8_t>());
    hipcub::CountingInputIterator<int> it(0);
    size_t b = 0, b2 = 0;
    CP_HIP(hipcub::DeviceSelect::Flagged(nullptr, b, it, flag.as<uint8_t>(), sel_q.as<int>(), nsel.as<int>(),
                                         (int)P, st));
    CP_HIP(hipcub::DeviceSelect::Flagged(nullptr, b2, key.as<unsigned long long>(), flag.as<uint8_t>(),
                                         sel_k.as<unsigned long long>(), nsel.as<int>() + 1, (int)P, st));
    tmp.alloc(ctx, std::max(b, b2));
    b = tmp.bytes();
    CP_HIP(hipcub::DeviceSelect::Flagged(tmp.get(), b, it, flag.as<uint8_t>(), sel_q.as<int>(), nsel.as<int>(),
                                         (int)P, st));
    b = tmp.bytes();
    CP_HIP(hipcub::DeviceSelect::Flagged(tmp.get(), b, key.as<unsigned long long>(), flag.as<uint8_t>(),
                                         sel_k.as<unsigned long long>(), nsel.as<int>() + 1, (int)P, st));
    int n = 0;
    CP_HIP(hipMemcpyAsync(&n, nsel.get(), 4, hipMemcpyDeviceToHost, st));
    CP_HIP(hipStreamSynchronize(st));
    d_idx.alloc(ctx, (size_t)std::max(n, 1) * 4);
    if (n == 0) return 0;
    if (interior_only) {   // column-major order is already the reference's
        CP_HIP(hipMemcpyAsync(d_idx.get(), sel_q.get(), (size_t)n * 4, hipMemcpyDeviceToDevice, st));
        return n;
    }
    DevBuf k2(ctx, (size_t)n * 8);
    b = 0;
    CP_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, b, sel_k.as<unsigned long long>(),
                                              k2.as<unsigned long long>(), sel_q.as<int>(), d_idx.as<int>(), n, 0,
                                              64, st));
    if (tmp.bytes() < b) tmp.alloc(ctx, b);
    b = tmp.bytes();
    CP_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.get(), b, sel_k.as<unsigned long long>(),
                                              k2.as<unsigned long long>(), sel_q.as<int>(), d_idx.as<int>(), n, 0,
                                              64, st));
    return n;
}

void matched_filter_device(cp_ctx* ctx, const float* d_image, const float* d_iv, int ny, int nx, const double* kernel,
                           int kr, int kc, int normalize_zerosum, double sigma, int border, cp_mf_result* out) {
    CP_ARG(d_image && kernel && out, "NULL argument");
    CP_ARG(ny > 0 && nx > 0, "image must be non-empty");
    CP_ARG(border == CP_BORDER_REPLICATE || border == CP_BORDER_ZERO, "unknown border mode; use :replicate or :zero");
    CP_ARG(kr > 0 && kc > 0 && kr % 2 == 1 && kc % 2 == 1, "kernel dimensions must be odd");
    hipStream_t st = ctx->stream;
    // 1. kernel normalisation (detection.jl:116-137)
    const size_t NK = (size_t)kr * kc;
    double sumP = 0, sumP2 = 0;
    for (size_t k = 0; k < NK; ++k) {
        sumP += kernel[k];
        sumP2 += kernel[k] * kernel[k];
    }
    std::vector<double> K(kernel, kernel + NK), K2(NK);
    double knorm = 1.0;
    if (normalize_zerosum) {
        const double denom = sumP2 - sumP * sumP / (double)NK;
        if (denom > 0) {
            for (size_t k = 0; k < NK; ++k) K[k] = (kernel[k] - sumP / (double)NK) / denom;
            knorm = std::sqrt(denom);
        }
    } else if (sumP2 > 0) {
        for (size_t k = 0; k < NK; ++k) K[k] = kernel[k] / sumP2;
        knorm = std::sqrt(sumP2);
    }
    for (size_t k = 0; k < NK; ++k) K2[k] = K[k] * K[k];
    // 2. correlations (:139-168)
    const int64_t P = (int64_t)ny * nx;
    const unsigned g = (unsigned)((P + BLOCK - 1) / BLOCK);
    DevBuf sm(ctx, (size_t)P * 8), sig(ctx, (size_t)P * 8), tmp, w, wv, num, den;
    if (d_iv) den.alloc(ctx, (size_t)P * 8);
    if (mf_lds_bytes<4>(kr, kc, d_iv != nullptr) <= CORR_LDS_MAX) {
        // one fused pass over LDS-staged float32 planes
        DevBuf dK(ctx, NK * 8), dK2(ctx, NK * 8);
        CP_HIP(hipMemcpyAsync(dK.get(), K.data(), NK * 8, hipMemcpyHostToDevice, st));
        CP_HIP(hipMemcpyAsync(dK2.get(), K2.data(), NK * 8, hipMemcpyHostToDevice, st));
        constexpr int RY = 4;
        dim3 gg((unsigned)((ny + CT * RY - 1) / (CT * RY)), (unsigned)((nx + CT - 1) / CT));
        if (d_iv)
            launch_n("k_mf2d", k_mf2d<RY, true>, gg, dim3(CT * CT), mf_lds_bytes<RY>(kr, kc, true), st, d_image, d_iv,
                     ny, nx, (const double*)dK.as<double>(), (const double*)dK2.as<double>(), kr, kc, border, knorm,
                     sm.as<double>(), den.as<double>(), sig.as<double>());
        else
            launch_n("k_mf2d", k_mf2d<RY, false>, gg, dim3(CT * CT), mf_lds_bytes<RY>(kr, kc, false), st, d_image,
                     (const float*)nullptr, ny, nx, (const double*)dK.as<double>(),
                     (const double*)dK2.as<double>(), kr, kc, border, knorm, sm.as<double>(), (double*)nullptr,
                     sig.as<double>());
        CP_HIP(hipStreamSynchronize(st));   // K, K2 are local
    } else {
        CorrPlan pk, pk2;
        plan_corr(ctx, K.data(), kr, kc, pk);
        run_corr<float>(ctx, pk, d_image, ny, nx, border, sm.as<double>(), tmp);
        if (d_iv) {
            plan_corr(ctx, K2.data(), kr, kc, pk2);
            w.alloc(ctx, (size_t)P * 8);
            wv.alloc(ctx, (size_t)P * 8);
            num.alloc(ctx, (size_t)P * 8);
            launch(k_mf_weights, dim3(g), dim3(BLOCK), 0, st, P, d_image, d_iv, w.as<double>(), wv.as<double>());
            run_corr<double>(ctx, pk, wv.as<double>(), ny, nx, border, num.as<double>(), tmp);
            run_corr<double>(ctx, pk2, w.as<double>(), ny, nx, border, den.as<double>(), tmp);
        }
        launch(k_mf_significance, dim3(g), dim3(BLOCK), 0, st, P, (const double*)num.as<double>(),
               (const double*)(d_iv ? den.as<double>() : nullptr), (const double*)sm.as<double>(), knorm,
               sig.as<double>());
    }
    // 3-4. peaks: interior local maxima of the significance at or above sigma
    DevBuf idx;
    const int n = select_maxima(ctx, sig.as<double>(), ny, nx, true, sigma, idx);
    out->n_peaks = n;
    out->kernel_norm = knorm;
    const int m = std::min(n, std::max(out->peaks_cap, 0));
    std::vector<D2HSeg> segs;
    DevBuf py, px, ps, pf;
    if (m > 0) {
        py.alloc(ctx, (size_t)m * 4);
        px.alloc(ctx, (size_t)m * 4);
        ps.alloc(ctx, (size_t)m * 8);
        pf.alloc(ctx, (size_t)m * 8);
        launch(k_peak_out, dim3(cdiv(m, BLOCK)), dim3(BLOCK), 0, st, m, (const int*)idx.as<int>(), ny,
               (const double*)sig.as<double>(), (const double*)sm.as<double>(),
               (const double*)(d_iv ? den.as<double>() : nullptr), knorm * knorm, py.as<int>(), px.as<int>(),
               ps.as<double>(), pf.as<double>());
        segs.push_back({out->peak_y, py.get(), (size_t)m * 4});
        segs.push_back({out->peak_x, px.get(), (size_t)m * 4});
        segs.push_back({out->peak_significance, ps.get(), (size_t)m * 8});
        segs.push_back({out->peak_flux, pf.get(), (size_t)m * 8});
    }
    if (out->smoothed) segs.push_back({out->smoothed, sm.get(), (size_t)P * 8});
    if (out->significance) segs.push_back({out->significance, sig.get(), (size_t)P * 8});
    if (out->smoothed_inv_var && d_iv) segs.push_back({out->smoothed_inv_var, den.get(), (size_t)P * 8});
    staged_d2h(ctx, st, segs);
    if (out->kernel) std::copy(K.begin(), K.end(), out->kernel);
}

}  // namespace
}  // namespace cp

using namespace cp;

extern "C" {

int cp_correlate_device(cp_ctx* ctx, const double* d_img, int32_t ny, int32_t nx, const double* kernel, int32_t kr,
                        int32_t kc, int32_t border, double* d_out) {
    if (!ctx) return CP_ESTATE;
    return guarded(ctx, [&] {
        CP_ARG(d_img && kernel && d_out && ny > 0 && nx > 0, "bad arguments");
        CP_ARG(border == CP_BORDER_REPLICATE || border == CP_BORDER_ZERO,
               "unknown border mode; use :replicate or :zero");
        CP_HIP(hipSetDevice(ctx->device));
        CorrPlan pl;
        plan_corr(ctx, kernel, kr, kc, pl);
        DevBuf tmp;
        run_corr<double>(ctx, pl, d_img, ny, nx, border, d_out, tmp);
        CP_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int cp_correlate(cp_ctx* ctx, const double* img, int32_t ny, int32_t nx, const double* kernel, int32_t kr,
                 int32_t kc, int32_t border, double* out) {
    if (!ctx) return CP_ESTATE;
    return guarded(ctx, [&] {
        CP_ARG(img && kernel && out && ny > 0 && nx > 0, "bad arguments");
        CP_ARG(border == CP_BORDER_REPLICATE || border == CP_BORDER_ZERO,
               "unknown border mode; use :replicate or :zero");
        CP_HIP(hipSetDevice(ctx->device));
        hipStream_t st = ctx->stream;
        const size_t P = (size_t)ny * nx;
        CorrPlan pl;
        plan_corr(ctx, kernel, kr, kc, pl);
        DevBuf din(ctx, P * 8), dout(ctx, P * 8), tmp;
        CP_HIP(hipMemcpyAsync(din.get(), img, P * 8, hipMemcpyHostToDevice, st));
        run_corr<double>(ctx, pl, din.as<double>(), ny, nx, border, dout.as<double>(), tmp);
        staged_d2h(ctx, st, {D2HSeg{out, dout.get(), P * 8}});
    });
}

int cp_findlocalmaxima(cp_ctx* ctx, const double* img, int32_t ny, int32_t nx, int32_t edges, int32_t cap,
                       int32_t* peak_y, int32_t* peak_x, int32_t* n_peaks) {
    if (!ctx) return CP_ESTATE;
    return guarded(ctx, [&] {
        CP_ARG(img && n_peaks && ny > 0 && nx > 0 && cap >= 0, "bad arguments");
        *n_peaks = 0;
        if (ny < 3 && nx < 3) return;
        CP_ARG(!(edges && ((ny == 1 && nx >= 3) || (nx == 1 && ny >= 3))),
               "image too narrow for edge maxima (a 1-pixel-wide image of length >= 3)");
        CP_HIP(hipSetDevice(ctx->device));
        hipStream_t st = ctx->stream;
        const size_t P = (size_t)ny * nx;
        DevBuf din(ctx, P * 8), idx;
        CP_HIP(hipMemcpyAsync(din.get(), img, P * 8, hipMemcpyHostToDevice, st));
        const int n = select_maxima(ctx, din.as<double>(), ny, nx, !edges, -INFINITY, idx);
        *n_peaks = n;
        const int m = std::min(n, cap);
        if (m <= 0) return;
        DevBuf py(ctx, (size_t)m * 4), px(ctx, (size_t)m * 4);
        launch(k_peak_out, dim3(cdiv(m, BLOCK)), dim3(BLOCK), 0, st, m, (const int*)idx.as<int>(), ny,
               (const double*)nullptr, (const double*)nullptr, (const double*)nullptr, 0.0, py.as<int>(),
               px.as<int>(), (double*)nullptr, (double*)nullptr);
        staged_d2h(ctx, st, {D2HSeg{peak_y, py.get(), (size_t)m * 4}, D2HSeg{peak_x, px.get(), (size_t)m * 4}});
    });
}

int cp_matched_filter_device(cp_ctx* ctx, const float* d_image, const float* d_inv_var, int32_t ny, int32_t nx,
                             const double* kernel, int32_t kr, int32_t kc, int32_t normalize_zerosum, double sigma,
                             int32_t border, cp_mf_result* out) {
    if (!ctx) return CP_ESTATE;
    return guarded(ctx, [&] {
        CP_HIP(hipSetDevice(ctx->device));
        matched_filter_device(ctx, d_image, d_inv_var, ny, nx, kernel, kr, kc, normalize_zerosum, sigma, border,
                              out);
    });
}

int cp_matched_filter(cp_ctx* ctx, const float* image, const float* inv_var, int32_t ny, int32_t nx,
                      const double* kernel, int32_t kr, int32_t kc, int32_t normalize_zerosum, double sigma,
                      int32_t border, cp_mf_result* out) {
    if (!ctx) return CP_ESTATE;
    return guarded(ctx, [&] {
        CP_ARG(image && ny > 0 && nx > 0, "bad arguments");
        CP_HIP(hipSetDevice(ctx->device));
        hipStream_t st = ctx->stream;
        const size_t P = (size_t)ny * nx;
        DevBuf dimg(ctx, P * 4), div;
        CP_HIP(hipMemcpyAsync(dimg.get(), image, P * 4, hipMemcpyHostToDevice, st));
        if (inv_var) {
            div.alloc(ctx, P * 4);
            CP_HIP(hipMemcpyAsync(div.get(), inv_var, P * 4, hipMemcpyHostToDevice, st));
        }
        matched_filter_device(ctx, dimg.as<float>(), inv_var ? div.as<float>() : nullptr, ny, nx, kernel, kr, kc,
                              normalize_zerosum, sigma, border, out);
        CP_HIP(hipStreamSynchronize(st));
    });
}

}  // extern "C"
