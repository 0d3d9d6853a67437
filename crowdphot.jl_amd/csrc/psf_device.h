// psf_device.h -- device-side PSF model evaluation for gfx950.
//
// Each star is evaluated relative to its integer anchor (ay, ax): a pixel at
// offset (dy, dx) from the anchor has model coordinates (dy - fy, dx - fx)
// with fy = Y - ay, fx = X - ax formed once per star in float64.  This keeps
// float32 evaluation exact to ~1e-7 px on 8k images, where Y itself would
// lose 5e-4 px to float32 rounding.
//
// Formulas follow the reference models:
//   CircularGaussianPSF evaluate / evaluate_fg  src/psf/parametric_models.jl:77-113
//   CircularMoffatPSF   evaluate / evaluate_fg  src/psf/parametric_models.jl:693-724
//   ImagePSF bicubic (Keys a=-1/2)              src/psf/empirical_models.jl:185-318
//   GriddedPSFModel corner blend                src/psf/gridded_psf.jl:176-271 and the
//   specialised corner pass + chain rule        src/psf/psf_fitting.jl:1233-1321,
//                                                psf_photometry_simultaneous.jl:500-510
#pragma once

#include <hip/hip_runtime.h>

namespace cp {

enum : int { KIND_GAUSS = 0, KIND_MOFFAT = 1, KIND_EPSF = 2 };

// -4 ln 2 (GAUSS_PRE, parametric_models.jl:18)
#define CP_GAUSS_PRE (-2.772588722239781)

template <typename T>
struct ModelConst {
    T fwhm2, inv_norm_g;   // gauss: fwhm^2 and 1/norm with norm = -(pi fwhm^2 / GAUSS_PRE)
    T alpha2, beta, norm_m;  // moffat: alpha^2, beta, norm = pi alpha^2 / (beta - 1)
    T inv_fwhm2, inv_alpha2, inv_norm_m;   // their reciprocals (no division per pixel)
    // ePSF geometry (all nodes share it)
    int gny, gnx, n_nodes;
    T osx, osy, ox, oy, fill;
    const float* grids;    // device, n_nodes * gny * gnx, node-major, each column-major
    T bkg;                 // fixed background added by every star's model
    int ios_y, ios_x;      // integer oversampling factors
    int Qy, Qx;            // polyphase plane sides of the edge-padded grid:
                           // ceil((gny + 2 PAD) / os_y), ceil((gnx + 2 PAD) / os_x)
    const float* poly;     // device: node 0 already in the padded polyphase LDS layout (or null)
};

// Edge padding of the LDS copy of an ePSF grid: PAD samples on every side
// repeat the nearest edge sample, so the reference's clamped stencil
// (bicubic_interpolate, empirical_models.jl:253-292) becomes a plain 16-tap
// read for every in-bounds pixel.
constexpr int EPSF_PAD = 2;

// Per-star constants of the single-node ePSF fast path.  With integer
// oversampling every pixel of a star shares the fractional grid offset, so the
// Keys cubic-convolution weights (the expansion of cubic4,
// empirical_models.jl:185-196) are fixed per star: value = sum_ab wy[a] wx[b] g,
// d/dv uses dwy, d/du uses dwx.  ra/ca are the polyphase LDS addresses of the
// four stencil rows / columns at the anchor pixel in the edge-padded grid; a
// pixel at offset (dy, dx) adds dy + dx * Qy.  Pixels outside [dylo, dyhi] x
// [dxlo, dxhi] are outside the grid's closed range (value fill); ok = 0 when
// no pixel is inside.
// With os_y = os_x = 4, tap (a, b) at pixel offset (dy, dx) also reads word
// base + dy + dx * Qy + off(phase, a, b), where phase = (ly mod 4) * 4 +
// (lx mod 4) of floor(v), floor(u) at the anchor and off depends only on
// the phase and the plane sides (epsf_tap_off): the renders' fixed-layout
// path reads the 16 taps at constant offsets.
template <typename T>
struct alignas(16) EpsfRec {   // 16-byte aligned: stored and loaded as whole 16-byte words
    T wy[4], wx[4], dwy[4], dwx[4];
    int ra[4], ca[4];
    int dylo, dyhi, dxlo, dxhi, ok;
    int phase, base;
};

// Word offset of tap (a, b) from EpsfRec::base + dy + dx * Qy for a star of
// the given phase, os_y = os_x = 4, polyphase planes of Qy x Qx words.
__host__ __device__ constexpr int epsf_tap_off(int phase, int a, int b, int Qy, int Qx) {
    return ((((phase >> 2) + a) & 3) * 4 + (((phase & 3) + b) & 3)) * (Qy * Qx) + (((phase & 3) + b) >> 2) * Qy +
           (((phase >> 2) + a) >> 2);
}

// Per-star record used by the gather kernels (rebuilt whenever theta changes).
template <typename T>
struct StarRec {
    T fy, fx;     // Y - ay, X - ax
    T flux;
    T amp;        // gauss: flux / norm; moffat: flux (beta-1) / (pi alpha^2); ePSF: flux
    int ay, ax;
    int mR;       // model-box half-width
    int flags;    // bit0 live, bit1 frozen at the current iteration
};

// Bilinear corner weights of a multi-node ePSF (gridded_psf.jl:176-208).
template <typename T>
struct CornerRec {
    int idx[4];   // 0-based node ids; weight 0 slots point at node 0
    T w[4], dwdy[4], dwdx[4];
};

// Keys a = -1/2 weights of cubic4 (empirical_models.jl:185-196) and their t-derivatives.
template <typename T>
__host__ __device__ __forceinline__ void keys_weights(T t, T w[4], T dw[4]) {
    const T t2 = t * t, t3 = t2 * t;
    w[0] = T(-0.5) * t3 + t2 - T(0.5) * t;
    w[1] = T(1.5) * t3 - T(2.5) * t2 + T(1);
    w[2] = T(-1.5) * t3 + T(2) * t2 + T(0.5) * t;
    w[3] = T(0.5) * t3 - T(0.5) * t2;
    dw[0] = T(-1.5) * t2 + T(2) * t - T(0.5);
    dw[1] = T(4.5) * t2 - T(5) * t;
    dw[2] = T(-4.5) * t2 + T(4) * t + T(0.5);
    dw[3] = T(1.5) * t2 - t;
}

// Polyphase address of 0-based grid sample (r, c): plane (r mod osy, c mod osx),
// column-major inside the plane.  Consecutive pixels of one star then read
// consecutive LDS words.
__host__ __device__ __forceinline__ int poly_addr(int r, int c, int osy, int osx, int Qy, int Qx) {
    return ((r % osy) * osx + (c % osx)) * (Qy * Qx) + (c / osx) * Qy + (r / osy);
}

// Stage one node grid (global, column-major) into LDS in the polyphase layout,
// edge-padded by EPSF_PAD samples (padded sample (r, c) holds grid sample
// (clamp(r - PAD), clamp(c - PAD))).
__device__ __forceinline__ void stage_poly(const float* __restrict__ g, int gny, int gnx, int osy, int osx,
                                           int Qy, int Qx, float* __restrict__ sg, int tid, int nthr) {
    const int py = gny + 2 * EPSF_PAD, px = gnx + 2 * EPSF_PAD;
    for (int i = tid; i < py * px; i += nthr) {
        const int c = i / py, r = i - c * py;
        const int rs = min(max(r - EPSF_PAD, 0), gny - 1), cs = min(max(c - EPSF_PAD, 0), gnx - 1);
        sg[poly_addr(r, c, osy, osx, Qy, Qx)] = g[rs + (size_t)cs * gny];
    }
}

// Copy the pre-staged grid (global, already in the LDS layout) into LDS.
__device__ __forceinline__ void load_poly(const float* __restrict__ src, float* __restrict__ dst, int n, int tid,
                                          int nthr) {
    const int n4 = n >> 2;
    const float4* s4 = reinterpret_cast<const float4*>(src);
    float4* d4 = reinterpret_cast<float4*>(dst);
    for (int i = tid; i < n4; i += nthr) d4[i] = s4[i];
    for (int i = (n4 << 2) + tid; i < n; i += nthr) dst[i] = src[i];
}

// floor(a / b) and ceil(a / b) for b > 0
__host__ __device__ __forceinline__ int floor_div(int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }
__host__ __device__ __forceinline__ int ceil_div(int a, int b) { return -floor_div(-a, b); }

// Build the per-star fast-path constants from its fractional position.
template <typename T>
__device__ __forceinline__ void epsf_rec(const ModelConst<T>& M, double fy, double fx, EpsfRec<T>& e) {
    double v0 = (double)M.ios_y * (0.0 - fy) + (double)M.oy;   // v at the anchor pixel
    double u0 = (double)M.ios_x * (0.0 - fx) + (double)M.ox;
    if (!(fabs(v0) < 1e8 && fabs(u0) < 1e8)) {  // non-finite or absurd position: every pixel off-grid
        v0 = -1e8;
        u0 = -1e8;
    }
    const double fly = floor(v0), flx = floor(u0);
    const int ly = (int)fly, lx = (int)flx;   // floor(v), floor(u) at the anchor pixel (1-based grid coordinates)
    T wd[4];
    keys_weights<T>((T)(v0 - fly), e.wy, wd);
    for (int a = 0; a < 4; ++a) e.dwy[a] = wd[a];
    keys_weights<T>((T)(u0 - flx), e.wx, wd);
    for (int a = 0; a < 4; ++a) e.dwx[a] = wd[a];
    // pixel offsets whose floor(v) is inside the closed range: [1, gny - 1],
    // or gny exactly when v = gny (t = 0, wy[1] = 1)
    const int osy = M.ios_y, osx = M.ios_x;
    e.dylo = ceil_div(1 - ly, osy);
    e.dyhi = (e.wy[1] == T(1) && floor_div(M.gny - ly, osy) * osy == M.gny - ly) ? floor_div(M.gny - ly, osy)
                                                                                 : floor_div(M.gny - 1 - ly, osy);
    e.dxlo = ceil_div(1 - lx, osx);
    e.dxhi = (e.wx[1] == T(1) && floor_div(M.gnx - lx, osx) * osx == M.gnx - lx) ? floor_div(M.gnx - lx, osx)
                                                                                 : floor_div(M.gnx - 1 - lx, osx);
    e.ok = e.dylo <= e.dyhi && e.dxlo <= e.dxhi;
    const int QQ = M.Qy * M.Qx;
    for (int a = 0; a < 4; ++a) {
        // stencil entry a reads 0-based grid row ly - 2 + a, padded row ly + a;
        // addresses are those of the in-range pixel (dylo, dxlo) shifted back
        // by (dylo, dxlo), so every clamped pixel offset lands in the grid
        const int r = e.ok ? ly - 2 + a + EPSF_PAD + osy * e.dylo : 0;
        const int c = e.ok ? lx - 2 + a + EPSF_PAD + osx * e.dxlo : 0;
        e.ra[a] = (r % osy) * osx * QQ + r / osy - (e.ok ? e.dylo : 0);
        e.ca[a] = (c % osx) * QQ + (c / osx - (e.ok ? e.dxlo : 0)) * M.Qy;
    }
    if (!e.ok) e.dylo = e.dyhi = e.dxlo = e.dxhi = 0;
    // padded row ly + a + 4 dy = 4 (Ly + dy) + (ly mod 4) + a, likewise columns
    const int Ly = floor_div(ly, 4), Lx = floor_div(lx, 4);
    e.phase = (ly - 4 * Ly) * 4 + (lx - 4 * Lx);
    e.base = Ly + Lx * M.Qy;
}

// ePSF value (no flux / bkg) of a star at pixel offset (dy, dx) and its
// derivatives along v and u, from the edge-padded polyphase LDS grid: the
// offset is clamped to the star's in-range window (so the 16 reads stay in
// the grid), and pixels outside the window take the fill value.
template <typename T, bool DERIV>
__device__ __forceinline__ void epsf_eval(const ModelConst<T>& M, const EpsfRec<T>& e, const float* __restrict__ sg,
                                          int dy, int dx, T& val, T& dpdv, T& dpdu) {
    const bool inb = e.ok && dy >= e.dylo && dy <= e.dyhi && dx >= e.dxlo && dx <= e.dxhi;
    const int dyc = min(max(dy, e.dylo), e.dyhi), dxc = min(max(dx, e.dxlo), e.dxhi);
    const float* sp = sg + (dyc + dxc * M.Qy);
    float g[16];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) g[a * 4 + b] = sp[e.ra[a] + e.ca[b]];
    T row[4], drow[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        row[a] = e.wx[0] * (T)g[a * 4] + e.wx[1] * (T)g[a * 4 + 1] + e.wx[2] * (T)g[a * 4 + 2] + e.wx[3] * (T)g[a * 4 + 3];
        if (DERIV)
            drow[a] = e.dwx[0] * (T)g[a * 4] + e.dwx[1] * (T)g[a * 4 + 1] + e.dwx[2] * (T)g[a * 4 + 2] +
                      e.dwx[3] * (T)g[a * 4 + 3];
    }
    val = e.wy[0] * row[0] + e.wy[1] * row[1] + e.wy[2] * row[2] + e.wy[3] * row[3];
    if (DERIV) {
        dpdv = e.dwy[0] * row[0] + e.dwy[1] * row[1] + e.dwy[2] * row[2] + e.dwy[3] * row[3];
        dpdu = e.wy[0] * drow[0] + e.wy[1] * drow[1] + e.wy[2] * drow[2] + e.wy[3] * drow[3];
    }
    if (!inb) {
        val = M.fill;
        dpdv = dpdu = T(0);
    }
}

template <typename T>
__device__ __forceinline__ void cubic4(T f1, T f2, T f3, T f4, T t, T& val, T& der) {
    T c1 = (f3 - f1) / T(2);
    T c4 = f3 - f2 - c1;
    T c2 = T(3) * c4 - (f4 - f2) / T(2) + c1;
    T c3 = c4 - c2;
    T c4t = t * c3;
    val = t * (t * (c4t + c2) + c1) + f2;
    der = t * (T(3) * c4t + T(2) * c2) + c1;
}

template <typename T>
__device__ __forceinline__ T cubic4v(T f1, T f2, T f3, T f4, T t) {
    T c1 = (f3 - f1) / T(2);
    T c4 = f3 - f2 - c1;
    T c2 = T(3) * c4 - (f4 - f2) / T(2) + c1;
    T c3 = c4 - c2;
    return t * (t * (t * c3 + c2) + c1) + f2;
}

// Stencil location of a bicubic evaluation (bicubic_interpolate, :239-292).
template <typename T>
struct Stencil {
    bool inb;
    T dx, dy;
    int ix[4], iy[4];  // 0-based, clamped
};

template <typename T>
__device__ __forceinline__ Stencil<T> stencil(int ny, int nx, T y, T x) {
    Stencil<T> s;
    s.inb = isfinite(x) && isfinite(y) && x >= T(1) && x <= T(nx) && y >= T(1) && y <= T(ny);
    T xc = s.inb ? x : T(1), yc = s.inb ? y : T(1);
    int lx = min(max((int)floor(xc), 1), nx - 1);
    int ly = min(max((int)floor(yc), 1), ny - 1);
    s.dx = xc - T(lx);
    s.dy = yc - T(ly);
    s.ix[0] = max(lx - 1, 1) - 1;
    s.ix[1] = lx - 1;
    s.ix[2] = lx;
    s.ix[3] = min(lx + 2, nx) - 1;
    s.iy[0] = max(ly - 1, 1) - 1;
    s.iy[1] = ly - 1;
    s.iy[2] = ly;
    s.iy[3] = min(ly + 2, ny) - 1;
    return s;
}

// Value only.
template <typename T>
__device__ __forceinline__ T bicubic_value(const float* __restrict__ g, int ny, int nx, T v, T u,
                                           T fill) {
    Stencil<T> s = stencil<T>(ny, nx, v, u);
    T row[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const float* col = g + s.iy[r];
        row[r] = cubic4v<T>((T)col[s.ix[0] * ny], (T)col[s.ix[1] * ny], (T)col[s.ix[2] * ny],
                            (T)col[s.ix[3] * ny], s.dx);
    }
    T val = cubic4v<T>(row[0], row[1], row[2], row[3], s.dy);
    return s.inb ? val : fill;
}

// Value and grid-space derivatives (value, d/dv, d/du).
template <typename T>
__device__ __forceinline__ void bicubic_vdd(const float* __restrict__ g, int ny, int nx, T v, T u,
                                            T fill, T& val, T& dpdv, T& dpdu) {
    Stencil<T> s = stencil<T>(ny, nx, v, u);
    T row[4], drow[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const float* col = g + s.iy[r];
        cubic4<T>((T)col[s.ix[0] * ny], (T)col[s.ix[1] * ny], (T)col[s.ix[2] * ny],
                  (T)col[s.ix[3] * ny], s.dx, row[r], drow[r]);
    }
    T vv, dv, du, unused;
    cubic4<T>(row[0], row[1], row[2], row[3], s.dy, vv, dv);
    cubic4<T>(drow[0], drow[1], drow[2], drow[3], s.dy, du, unused);
    val = s.inb ? vv : fill;
    dpdv = s.inb ? dv : T(0);
    dpdu = s.inb ? du : T(0);
}

// Model value of star r at model-frame offset (ddy, ddx) = (py - Y, px - X)
// (background included, as evaluate does).
template <typename T, int KIND>
__device__ __forceinline__ T star_value(const ModelConst<T>& M, const StarRec<T>& r, T ddy, T ddx,
                                        const float* __restrict__ grid,
                                        const CornerRec<T>* __restrict__ cr) {
    if constexpr (KIND == KIND_GAUSS) {
        T sq = (ddx * ddx + ddy * ddy) * M.inv_fwhm2;
        return fma(r.amp, exp(T(CP_GAUSS_PRE) * sq), M.bkg);
    } else if constexpr (KIND == KIND_MOFFAT) {
        T r2 = ddx * ddx + ddy * ddy;
        return fma(r.amp, pow(fma(r2, M.inv_alpha2, T(1)), -M.beta), M.bkg);
    } else {
        T u = M.osx * ddx + M.ox;
        T v = M.osy * ddy + M.oy;
        if (M.n_nodes == 1) {
            T p = bicubic_value<T>(grid, M.gny, M.gnx, v, u, M.fill);
            return fma(r.flux, p, M.bkg);
        }
        T s = T(0);
        const size_t nn = (size_t)M.gny * M.gnx;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            if (cr->w[c] == T(0)) continue;
            T p = bicubic_value<T>(M.grids + cr->idx[c] * nn, M.gny, M.gnx, v, u, M.fill);
            s = fma(cr->w[c], p, s);
        }
        return fma(r.flux, s, M.bkg);
    }
}

// Gradient entries (d/dy, d/dx, d/dflux) of star r at model-frame offset (ddy, ddx).
template <typename T, int KIND>
__device__ __forceinline__ void star_grad(const ModelConst<T>& M, const StarRec<T>& r, T ddy, T ddx,
                                          const float* __restrict__ grid,
                                          const CornerRec<T>* __restrict__ cr, T& gy, T& gx, T& gf) {
    if constexpr (KIND == KIND_GAUSS) {
        T sq = (ddx * ddx + ddy * ddy) * M.inv_fwhm2;
        T g = exp(T(CP_GAUSS_PRE) * sq);
        T Ag = r.amp * g;
        T gf2 = T(CP_GAUSS_PRE) * M.inv_fwhm2;
        gy = T(-2) * Ag * gf2 * ddy;
        gx = T(-2) * Ag * gf2 * ddx;
        gf = g * M.inv_norm_g;
    } else if constexpr (KIND == KIND_MOFFAT) {
        T r2 = ddx * ddx + ddy * ddy;
        T u = fma(r2, M.inv_alpha2, T(1));
        T prof = pow(u, -M.beta);
        T amp = r.flux * M.inv_norm_m;
        T Ag = amp * prof;
        T c = T(2) * Ag * M.beta * M.inv_alpha2 / u;   // 2 A g beta / (alpha^2 u)
        gy = c * ddy;
        gx = c * ddx;
        gf = prof * M.inv_norm_m;
    } else {
        T u = M.osx * ddx + M.ox;
        T v = M.osy * ddy + M.oy;
        if (M.n_nodes == 1) {
            T pv, pdv, pdu;
            bicubic_vdd<T>(grid, M.gny, M.gnx, v, u, M.fill, pv, pdv, pdu);
            gy = r.flux * (-(M.osy * pdv));
            gx = r.flux * (-(M.osx * pdu));
            gf = pv;
            return;
        }
        T s = T(0), dsdy = T(0), dsdx = T(0);
        const size_t nn = (size_t)M.gny * M.gnx;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            if (cr->w[c] == T(0)) continue;
            T pv, pdv, pdu;
            bicubic_vdd<T>(M.grids + cr->idx[c] * nn, M.gny, M.gnx, v, u, M.fill, pv, pdv, pdu);
            s = fma(cr->w[c], pv, s);
            dsdy += cr->dwdy[c] * pv - cr->w[c] * M.osy * pdv;
            dsdx += cr->dwdx[c] * pv - cr->w[c] * M.osx * pdu;
        }
        gy = r.flux * dsdy;
        gx = r.flux * dsdx;
        gf = s;
    }
}

}  // namespace cp
