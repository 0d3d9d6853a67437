// normal_op.h -- the inner least-squares solve of the simultaneous fit in
// star space (psf_photometry_simultaneous.jl:1164,1221-1224 over the operator
// of :55-154).
//
// Within one outer iteration the equilibrated stamp Jacobian J (the V rows)
// is fixed, so its normal matrix H = J^T J is too: a sparse matrix of p x p
// blocks H_ab = sum over the pixels two stamps share of V_a V_b^T, one per
// pair of stars whose (2R+1)^2 stamps overlap (anchors within 2R on both
// axes).  It is assembled once per outer iteration from V, in float64, and
// every Golub-Kahan step of LSQR / LSMR and every CG step then needs only
// H v, in star space:
//
//   J^T u_{k+1} = (H v_k - alpha_k J^T u_k) / beta_{k+1},
//   beta_{k+1}^2 = v_k^T H v_k - 2 alpha_k v_k^T J^T u_k + alpha_k^2,
//   t_{k+1} = J^T u_{k+1} - beta_{k+1} v_k,  alpha_{k+1} = |t_{k+1}|,
//
// which is the bidiagonalization u_{k+1} beta_{k+1} = J v_k - alpha_k u_k,
// v_{k+1} alpha_{k+1} = J^T u_{k+1} - beta_{k+1} v_k with the pixel vector u
// never formed (|u_k| = 1).  The scalar recurrences and the x / w (LSMR:
// h, hbar) updates are those of the pixel-space method, unchanged.  Per
// iteration the traffic is the blocks and a few star vectors instead of two
// passes over every stamp row.
//
// Neighbour lists: stars are binned into cells of side 2R (anchors shifted by
// R - 1 so every active anchor is inside the grid), so the stars whose stamps
// overlap a star's are in its cell or the 8 around it; each star's row lists
// them in cell order, Morton order inside a cell.
#pragma once

#include "sim_kernels.h"

namespace cp {

constexpr int HB = 12;   // doubles per block: p <= 3 rows of 4 (the 4th column 0)

struct CellGeo {
    int CS, ncy, ncx, R;
};

__device__ __forceinline__ int cell_of(const CellGeo& g, int ay, int ax) {
    const int cy = (ay + g.R - 1) / g.CS, cx = (ax + g.R - 1) / g.CS;
    return cx * g.ncy + cy;
}

static __global__ void k_cell_keys(int na, const int* __restrict__ ay, const int* __restrict__ ax, CellGeo g,
                                   unsigned int* __restrict__ keys, int* __restrict__ vals) {
    const int a = blockIdx.x * BLOCK + threadIdx.x;
    if (a >= na) return;
    keys[a] = (unsigned int)cell_of(g, ay[a], ax[a]);
    vals[a] = a;
}

// Pairs (a, b) of stars whose stamps overlap; count (emit = false) or write
// the row of star a (emit = true) at rowptr[a].
template <bool EMIT>
static __global__ void k_pairs(int na, const int* __restrict__ ay, const int* __restrict__ ax, CellGeo g,
                               const int* __restrict__ cstart, const int* __restrict__ cend,
                               const int* __restrict__ clist, int* __restrict__ cnt,
                               const int* __restrict__ rowptr, int* __restrict__ pb) {
    const int a = blockIdx.x * BLOCK + threadIdx.x;
    if (a >= na) return;
    const int ya = ay[a], xa = ax[a], D = 2 * g.R;
    const int cy = (ya + g.R - 1) / g.CS, cx = (xa + g.R - 1) / g.CS;
    int o = EMIT ? rowptr[a] : 0;
    for (int jx = max(cx - 1, 0); jx <= min(cx + 1, g.ncx - 1); ++jx)
        for (int jy = max(cy - 1, 0); jy <= min(cy + 1, g.ncy - 1); ++jy) {
            const int c = jx * g.ncy + jy;
            for (int i = cstart[c]; i < cend[c]; ++i) {
                const int b = clist[i];
                if (abs(ay[b] - ya) <= D && abs(ax[b] - xa) <= D) {
                    if (EMIT) pb[o] = b;
                    ++o;
                }
            }
        }
    if (!EMIT) cnt[a] = o;
}

// H_ab for every pair of the live rows in lst (Sc->n_jtu_stars of them):
// GL lanes per row, its pairs in turn, lanes striding over the pixels the two
// stamps share; the nine sums are group reductions in float64.  Blocks of a
// frozen neighbour are written as zeros (its vectors are not read).
template <typename T, int GL>
__global__ __launch_bounds__(BLOCK) void k_hbuild(const int* __restrict__ lst, const uint8_t* __restrict__ live,
                                                  const int* __restrict__ rowptr, const int* __restrict__ pb,
                                                  const int* __restrict__ ay, const int* __restrict__ ax, int R,
                                                  int S, const T* __restrict__ V, double* __restrict__ H,
                                                  const Scal* Sc) {
    constexpr int GPB = BLOCK / GL;
    const int sub = threadIdx.x % GL;
    const int S2 = S * S;
    const int nl = Sc->n_jtu_stars;
    const vec4_t<T>* V4 = reinterpret_cast<const vec4_t<T>*>(V);
    for (int li = xcd_remap(blockIdx.x, gridDim.x) * GPB + threadIdx.x / GL; li < nl; li += gridDim.x * GPB) {
        const int a = lst[li];
        if (!live[a]) continue;   // uniform within the group
        const int ya = ay[a], xa = ax[a];
        for (int j = rowptr[a]; j < rowptr[a + 1]; ++j) {
            const int b = pb[j];
            double h[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            if (live[b]) {
                const int dy0 = ay[b] - ya, dx0 = ax[b] - xa;
                const int ylo = max(-R, dy0 - R), yhi = min(R, dy0 + R);
                const int xlo = max(-R, dx0 - R), xhi = min(R, dx0 + R);
                const int hy = yhi - ylo + 1, nov = hy * (xhi - xlo + 1);
                for (int i = sub; i < nov; i += GL) {
                    const int dy = ylo + i % hy, dx = xlo + i / hy;
                    const vec4_t<T> va = V4[(size_t)a * S2 + (dx + R) * S + (dy + R)];
                    const vec4_t<T> vb = V4[(size_t)b * S2 + (dx - dx0 + R) * S + (dy - dy0 + R)];
                    const double A[3] = {(double)va.x, (double)va.y, (double)va.z};
                    const double B[3] = {(double)vb.x, (double)vb.y, (double)vb.z};
#pragma unroll
                    for (int k = 0; k < 3; ++k)
#pragma unroll
                        for (int l = 0; l < 3; ++l) h[3 * k + l] += A[k] * B[l];
                }
#pragma unroll
                for (int e = 0; e < 9; ++e) h[e] = group_sum<GL>(h[e]);
            }
            // lane e < 12 writes slot e of the block (row k = e / 4, column e % 4)
            if (sub < HB) {
                const int k = sub >> 2, l = sub & 3;
                double v = 0;
#pragma unroll
                for (int e = 0; e < 9; ++e)
                    if (e == 3 * k + l && l < 3) v = h[e];
                H[(size_t)j * HB + sub] = v;
            }
        }
    }
}

__device__ __forceinline__ double4 ld4(const double* p) { return *reinterpret_cast<const double4*>(p); }

// y_a = sum_b H_ab v_b for the rows in lst, one thread per row, with the
// partial sums the next scalar step needs.
//  * LSQR / LSMR: v = t / alpha (the first call: t_1 = -b_scaled / beta1);
//    partials v . y and v . J^T u_k.
//  * CG: the direction p = s + beta p_old of each row is formed where it is
//    gathered (the first call: p_0 = s_0 = -b_scaled) and each row writes its
//    own to pnew; partials p . H p and |p|^2.
// Neighbours that are not live contribute nothing (their vectors may be stale).
template <int SOLVER>
__global__ __launch_bounds__(BLOCK) void k_hv(const int* __restrict__ lst, const uint8_t* __restrict__ live,
                                              const int* __restrict__ rowptr, const int* __restrict__ pb,
                                              const double* __restrict__ H, const double* __restrict__ b_scaled,
                                              const double* __restrict__ tbuf, const double* __restrict__ ju,
                                              const double* __restrict__ pold, double* __restrict__ pnew,
                                              double* __restrict__ y, int first, double* __restrict__ partials,
                                              double* __restrict__ partials2, Scal* Sc) {
    if (Sc->done) return;
    __shared__ double sred[NWAVE];
    constexpr bool CG = SOLVER == CP_SOLVER_CGLS;
    const double alpha = Sc->alpha;
    const double ia = alpha != 0 ? 1.0 / alpha : 1.0;
    const double ib1 = (!CG && first) ? 1.0 / Sc->beta1 : 0.0;
    const double fac = CG ? 1.0 : (first ? -ib1 * ia : ia);
    const double cg_b = CG && !first ? Sc->cg_beta : 0.0;
    const double* src = first ? b_scaled : tbuf;
    const double sgn = (CG && first) ? -1.0 : 1.0;
    const int nl = Sc->n_jtu_stars;
    double s1 = 0, s2 = 0;
    for (int li = xcd_remap(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x; li < nl; li += gridDim.x * BLOCK) {
        const int a = lst[li];
        double4 acc = make_double4(0, 0, 0, 0);
        if (live[a]) {
            for (int j = rowptr[a]; j < rowptr[a + 1]; ++j) {
                const int b = pb[j];
                if (!live[b]) continue;
                double4 v = ld4(src + (size_t)b * VP);
                if (CG) {
                    if (first) {
                        v.x = -v.x; v.y = -v.y; v.z = -v.z;
                    } else {
                        const double4 po = ld4(pold + (size_t)b * VP);
                        v.x += cg_b * po.x; v.y += cg_b * po.y; v.z += cg_b * po.z;
                    }
                } else {
                    v.x *= fac; v.y *= fac; v.z *= fac;
                }
                const double* h = H + (size_t)j * HB;
                const double4 h0 = ld4(h), h1 = ld4(h + 4), h2 = ld4(h + 8);
                acc.x += h0.x * v.x + h0.y * v.y + h0.z * v.z;
                acc.y += h1.x * v.x + h1.y * v.y + h1.z * v.z;
                acc.z += h2.x * v.x + h2.y * v.y + h2.z * v.z;
            }
            double4 va = ld4(src + (size_t)a * VP);
            if (CG) {
                if (first) {
                    va.x = -va.x; va.y = -va.y; va.z = -va.z;
                } else {
                    const double4 po = ld4(pold + (size_t)a * VP);
                    va.x += cg_b * po.x; va.y += cg_b * po.y; va.z += cg_b * po.z;
                }
                *reinterpret_cast<double4*>(pnew + (size_t)a * VP) = make_double4(va.x, va.y, va.z, 0.0);
                s1 += va.x * acc.x + va.y * acc.y + va.z * acc.z;
                s2 += va.x * va.x + va.y * va.y + va.z * va.z;
            } else {
                va.x *= fac; va.y *= fac; va.z *= fac;
                double4 wk;
                if (first) {
                    wk = ld4(b_scaled + (size_t)a * VP);
                    wk.x *= -ib1; wk.y *= -ib1; wk.z *= -ib1;
                } else {
                    wk = ld4(ju + (size_t)a * VP);
                }
                s1 += va.x * acc.x + va.y * acc.y + va.z * acc.z;
                s2 += va.x * wk.x + va.y * wk.y + va.z * wk.z;
            }
        } else if (CG) {
            *reinterpret_cast<double4*>(pnew + (size_t)a * VP) = make_double4(0, 0, 0, 0);
        }
        (void)sgn;
        *reinterpret_cast<double4*>(y + (size_t)a * VP) = acc;
    }
    block_store(s1, partials, sred);
    block_store(s2, partials2, sred);
}

// Per star of lst (one thread each), after the scalars of this step:
//  * LSQR / LSMR: g = y - alpha J^T u_k is beta J^T u_{k+1}; J^T u_{k+1} =
//    g / beta and t <- J^T u_{k+1} - beta v, with the pending update of the
//    previous iteration folded in (hbar <- w - del hbar, x <- x + sig hbar,
//    w <- v - tau w; LSQR is del = 0).  Partials |t|^2, |w|^2 and, for LSMR,
//    |x|^2, x.w, x.hbar, w.hbar, |hbar|^2.
//  * CG: x <- x + a p, s <- s - a (H p + lam^2 p).  Partials |s|^2 and s . p.
template <int SOLVER>
__global__ __launch_bounds__(BLOCK) void k_gk_update(
    const int* __restrict__ lst, const uint8_t* __restrict__ live, int p, const double* __restrict__ y,
    const double* __restrict__ b_scaled, double* __restrict__ tbuf, double* __restrict__ ju,
    const double* __restrict__ pcur, double* __restrict__ x, double* __restrict__ wv, double* __restrict__ hb,
    int first, double* __restrict__ partials, double* __restrict__ partials2, double* __restrict__ partials3,
    Scal* Sc) {
    if (Sc->done) return;
    __shared__ double sred[NWAVE];
    constexpr bool CG = SOLVER == CP_SOLVER_CGLS, LSMR = SOLVER == CP_SOLVER_LSMR;
    const double alpha = Sc->alpha, beta = Sc->beta_next;
    const double ia = alpha != 0 ? 1.0 / alpha : 1.0;
    const double ib = beta != 0 ? 1.0 / beta : 1.0;
    const double sig = first ? 0.0 : Sc->sig, tau = first ? 0.0 : Sc->tau, del = first ? 0.0 : Sc->del;
    const double ib1 = (!CG && first) ? 1.0 / Sc->beta1 : 0.0;
    const double cg_a = CG ? Sc->cg_alpha : 0.0;
    const double lam2 = CG ? Sc->lam * Sc->lam : 0.0;
    double tsq = 0, wsq = 0, xx = 0, xw = 0, xh = 0, wh = 0, hh = 0;
    const int nl = Sc->n_jtu_stars;
    for (int li = xcd_remap(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x; li < nl; li += gridDim.x * BLOCK) {
        const int a = lst[li];
        const size_t base = (size_t)a * p, b4 = (size_t)a * VP;
        if (!live[a]) {
            if (first)
                for (int k = 0; k < p; ++k) {
                    x[base + k] = 0;
                    wv[base + k] = 0;
                    if (LSMR) hb[base + k] = 0;
                    tbuf[b4 + k] = 0;
                    ju[b4 + k] = 0;
                }
            continue;
        }
        for (int k = 0; k < p; ++k) {
            const double yk = y[b4 + k];
            if (CG) {
                const double pk = pcur[b4 + k];
                const double zk = yk + lam2 * pk;
                x[base + k] = (first ? 0.0 : x[base + k]) + cg_a * pk;
                const double sk = first ? -b_scaled[b4 + k] : tbuf[b4 + k];
                const double sn = sk - cg_a * zk;
                tbuf[b4 + k] = sn;
                tsq += sn * sn;
                wsq += sn * pk;
            } else {
                const double told = first ? -b_scaled[b4 + k] * ib1 : tbuf[b4 + k];
                const double v = told * ia;
                const double wk = first ? told : ju[b4 + k];   // J^T u_k (u_1 = -r / beta1)
                const double g = yk - alpha * wk;
                const double xo = first ? 0.0 : x[base + k];
                const double wo = first ? 0.0 : wv[base + k];
                const double ho = (LSMR && !first) ? hb[base + k] : 0.0;
                const double hn = wo - del * ho;
                const double xn = xo + sig * hn;
                x[base + k] = xn;
                const double wn = v - tau * wo;
                wv[base + k] = wn;
                if (LSMR) {
                    hb[base + k] = hn;
                    xx += xn * xn;
                    xw += xn * wn;
                    xh += xn * hn;
                    wh += wn * hn;
                    hh += hn * hn;
                }
                wsq += wn * wn;
                const double jn = beta != 0 ? g * ib : wk;
                ju[b4 + k] = jn;
                const double tn = beta != 0 ? jn - beta * v : told;
                tbuf[b4 + k] = tn;
                tsq += tn * tn;
            }
        }
    }
    block_store(tsq, partials, sred);
    block_store(wsq, partials2, sred);
    if (LSMR) {
        const int nb = gridDim.x;
        block_store(xx, partials3, sred);
        block_store(xw, partials3 + nb, sred);
        block_store(xh, partials3 + 2 * nb, sred);
        block_store(wh, partials3 + 3 * nb, sred);
        block_store(hh, partials3 + 4 * nb, sred);
    }
}

}  // namespace cp
