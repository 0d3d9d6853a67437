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
    // ePSF geometry (all nodes share it)
    int gny, gnx, n_nodes;
    T osx, osy, ox, oy, fill;
    const float* grids;    // device, n_nodes * gny * gnx, node-major, each column-major
    T bkg;                 // fixed background added by every star's model
    int ios_y, ios_x;      // integer oversampling factors
    int Qy, Qx;            // polyphase plane sides: ceil(gny / os_y), ceil(gnx / os_x)
};

// Per-star constants of the single-node ePSF fast path.  With integer
// oversampling every pixel of a star shares the fractional grid offset, so the
// Keys cubic-convolution weights (the expansion of cubic4,
// empirical_models.jl:185-196) are fixed per star: value = sum_ab wy[a] wx[b] g,
// d/dv uses dwy, d/du uses dwx.  ra/ca are the polyphase LDS addresses of the
// four stencil rows / columns at the anchor pixel; a pixel at offset (dy, dx)
// adds dy + dx * Qy.
template <typename T>
struct EpsfRec {
    T wy[4], wx[4], dwy[4], dwx[4];
    int ly, lx;      // floor(v), floor(u) at the anchor pixel (1-based grid coordinates)
    int ra[4], ca[4];
    int ih;          // every pixel with |dy|, |dx| <= ih evaluates on the unclamped interior stencil
};

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

// Stage one node grid (global, column-major) into LDS in the polyphase layout.
__device__ __forceinline__ void stage_poly(const float* __restrict__ g, int gny, int gnx, int osy, int osx,
                                           int Qy, int Qx, float* __restrict__ sg, int tid, int nthr) {
    for (int i = tid; i < gny * gnx; i += nthr) {
        const int c = i / gny, r = i - c * gny;
        sg[poly_addr(r, c, osy, osx, Qy, Qx)] = g[i];
    }
}

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
    e.ly = (int)fly;
    e.lx = (int)flx;
    T wd[4];
    keys_weights<T>((T)(v0 - fly), e.wy, wd);
    for (int a = 0; a < 4; ++a) e.dwy[a] = wd[a];
    keys_weights<T>((T)(u0 - flx), e.wx, wd);
    for (int a = 0; a < 4; ++a) e.dwx[a] = wd[a];
    const int QQ = M.Qy * M.Qx;
    for (int a = 0; a < 4; ++a) {
        // 0-based row of stencil entry a is ly - 2 + a; floor division keeps
        // the plane arithmetic exact for negative rows (used only off the fast path)
        int r = e.ly - 2 + a, c = e.lx - 2 + a;
        int pr = ((r % M.ios_y) + M.ios_y) % M.ios_y, qr = (r - pr) / M.ios_y;
        int pc = ((c % M.ios_x) + M.ios_x) % M.ios_x, qc = (c - pc) / M.ios_x;
        e.ra[a] = pr * M.ios_x * QQ + qr;
        e.ca[a] = pc * QQ + qc * M.Qy;
    }
    // interior stencil rows / columns are 2 .. gn-2; the pixel offset d moves
    // the stencil by os * d grid samples
    auto room = [](int lo, int hi, int os) { return lo < 0 || hi < 0 ? -1 : min(lo, hi) / os; };
    e.ih = min(room(e.ly - 2, M.gny - 2 - e.ly, M.ios_y), room(e.lx - 2, M.gnx - 2 - e.lx, M.ios_x));
}

// Grid sample with the reference's stencil clamping, polyphase LDS layout.
__device__ __forceinline__ float poly_sample(const float* __restrict__ sg, int r, int c, int gny, int gnx, int osy,
                                             int osx, int Qy, int Qx) {
    r = min(max(r, 0), gny - 1);
    c = min(max(c, 0), gnx - 1);
    return sg[poly_addr(r, c, osy, osx, Qy, Qx)];
}

// ePSF value (no flux/bkg) of a star at pixel offset (dy, dx), fast path.
// Interior evaluation (stencil rows / columns within 2 .. gn-2 for this
// pixel): 16 LDS taps, no bounds tests.
template <typename T, bool DERIV>
__device__ __forceinline__ void epsf_interior(const ModelConst<T>& M, const EpsfRec<T>& e, const float* __restrict__ sg,
                                              int dy, int dx, T& val, T& dpdv, T& dpdu) {
    T row[4], drow[4];
    const int off = dy + dx * M.Qy;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const float* rp = sg + e.ra[a] + off;
        const T g0 = (T)rp[e.ca[0]], g1 = (T)rp[e.ca[1]], g2 = (T)rp[e.ca[2]], g3 = (T)rp[e.ca[3]];
        row[a] = e.wx[0] * g0 + e.wx[1] * g1 + e.wx[2] * g2 + e.wx[3] * g3;
        if (DERIV) drow[a] = e.dwx[0] * g0 + e.dwx[1] * g1 + e.dwx[2] * g2 + e.dwx[3] * g3;
    }
    val = e.wy[0] * row[0] + e.wy[1] * row[1] + e.wy[2] * row[2] + e.wy[3] * row[3];
    if (DERIV) {
        dpdv = e.dwy[0] * row[0] + e.dwy[1] * row[1] + e.dwy[2] * row[2] + e.dwy[3] * row[3];
        dpdu = e.wy[0] * drow[0] + e.wy[1] * drow[1] + e.wy[2] * drow[2] + e.wy[3] * drow[3];
    }
}

template <typename T, bool DERIV>
__device__ __forceinline__ void epsf_eval(const ModelConst<T>& M, const EpsfRec<T>& e, const float* __restrict__ sg,
                                          int dy, int dx, T& val, T& dpdv, T& dpdu) {
    const int ly = e.ly + M.ios_y * dy, lx = e.lx + M.ios_x * dx;
    // closed-bounds test of bicubic_interpolate: floor(v) in [1, gny-1], or v == gny exactly
    const bool inb = ly >= 1 && lx >= 1 && (ly < M.gny || (ly == M.gny && e.wy[1] == T(1))) &&
                     (lx < M.gnx || (lx == M.gnx && e.wx[1] == T(1)));
    if (!inb) {
        val = M.fill;
        dpdv = dpdu = T(0);
        return;
    }
    T row[4], drow[4];
    if (ly >= 2 && ly <= M.gny - 2 && lx >= 2 && lx <= M.gnx - 2) {
        epsf_interior<T, DERIV>(M, e, sg, dy, dx, val, dpdv, dpdu);
        return;
    } else {
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            T g[4];
#pragma unroll
            for (int b = 0; b < 4; ++b)
                g[b] = (T)poly_sample(sg, ly - 2 + a, lx - 2 + b, M.gny, M.gnx, M.ios_y, M.ios_x, M.Qy, M.Qx);
            row[a] = e.wx[0] * g[0] + e.wx[1] * g[1] + e.wx[2] * g[2] + e.wx[3] * g[3];
            if (DERIV) drow[a] = e.dwx[0] * g[0] + e.dwx[1] * g[1] + e.dwx[2] * g[2] + e.dwx[3] * g[3];
        }
    }
    val = e.wy[0] * row[0] + e.wy[1] * row[1] + e.wy[2] * row[2] + e.wy[3] * row[3];
    if (DERIV) {
        dpdv = e.dwy[0] * row[0] + e.dwy[1] * row[1] + e.dwy[2] * row[2] + e.dwy[3] * row[3];
        dpdu = e.wy[0] * drow[0] + e.wy[1] * drow[1] + e.wy[2] * drow[2] + e.wy[3] * drow[3];
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
        T sq = (ddx * ddx + ddy * ddy) / M.fwhm2;
        return fma(r.amp, exp(T(CP_GAUSS_PRE) * sq), M.bkg);
    } else if constexpr (KIND == KIND_MOFFAT) {
        T r2 = ddx * ddx + ddy * ddy;
        return fma(r.amp, pow(T(1) + r2 / M.alpha2, -M.beta), M.bkg);
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
        T sq = (ddx * ddx + ddy * ddy) / M.fwhm2;
        T g = exp(T(CP_GAUSS_PRE) * sq);
        T Ag = r.amp * g;
        T gf2 = T(CP_GAUSS_PRE) / M.fwhm2;
        gy = T(-2) * Ag * gf2 * ddy;
        gx = T(-2) * Ag * gf2 * ddx;
        gf = g * M.inv_norm_g;
    } else if constexpr (KIND == KIND_MOFFAT) {
        T r2 = ddx * ddx + ddy * ddy;
        T u = T(1) + r2 / M.alpha2;
        T prof = pow(u, -M.beta);
        T amp = r.flux / M.norm_m;
        T Ag = amp * prof;
        gy = T(2) * Ag * M.beta * ddy / (M.alpha2 * u);
        gx = T(2) * Ag * M.beta * ddx / (M.alpha2 * u);
        gf = prof / M.norm_m;
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
