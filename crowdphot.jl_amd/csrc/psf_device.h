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
