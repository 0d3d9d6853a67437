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
// h, hbar) updates are those of the pixel-space method, unchanged.
//
// One launch per Golub-Kahan step (k_gk, one lane per row entry): each entry
// forms the v_k of its neighbour where it gathers it, from the g_{k-1} =
// H v_{k-1} - alpha_{k-1} J^T u_{k-1} and v_{k-1} the previous step stored,
//
//   v_k = (g_{k-1} / beta_k - beta_k v_{k-1}) / alpha_k,
//
// and the scalars need only sums over the rows of this step: with
// g_k = H v_k - alpha_k J^T u_k,
//
//   alpha_{k+1}^2 = |g_k|^2 / beta_{k+1}^2 - 2 g_k . v_k + beta_{k+1}^2 |v_k|^2,
//   g_k . v_k = v_k . H v_k - alpha_k v_k . J^T u_k,
//
// so one block-order reduction (k_finish_gk) per step gives beta_{k+1},
// alpha_{k+1} and the LSQR / LSMR rotations.  The expansion of alpha_{k+1}^2
// loses about eps (alpha^2 + beta^2) / alpha^2 relative, far below the
// tolerances the solve runs to.
//
// Storage: the blocks of one outer iteration live in the slots of the pair
// rows; the block of a pair (a, b) is built once, in the slot of the row of
// min(a, b) (its upper entry), and the other row reads it transposed.  Each
// row entry's (neighbour, code) word is rewritten with the blocks, so the
// operator kernels read no live flags.  Star vectors are float64 at stride p.
//
// Neighbour lists: stars are binned into cells of side 2R (anchors shifted by
// R - 1 so every active anchor is inside the grid), so the stars whose stamps
// overlap a star's are in its cell or the 8 around it; each star's row lists
// the star itself first, then the others in cell order, Morton order inside
// a cell.
#pragma once

#include <climits>

#include "sim_kernels.h"

namespace cp {

// Calls f with the number of free parameters per star (1..3) as a constant.
template <typename F>
static inline void with_p(int p, F&& f) {
    if (p == 1) f(std::integral_constant<int, 1>{});
    else if (p == 2) f(std::integral_constant<int, 2>{});
    else f(std::integral_constant<int, 3>{});
}

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
// the row of star a (emit = true) at rowptr[a]: a itself, then the others in
// cell order.
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
    // the row's own entry first: it heads the row in the Golub-Kahan step
    if (EMIT) pb[o] = a;
    ++o;
    for (int jx = max(cx - 1, 0); jx <= min(cx + 1, g.ncx - 1); ++jx)
        for (int jy = max(cy - 1, 0); jy <= min(cy + 1, g.ncy - 1); ++jy) {
            const int c = jx * g.ncy + jy;
            for (int i = cstart[c]; i < cend[c]; ++i) {
                const int b = clist[i];
                if (b != a && abs(ay[b] - ya) <= D && abs(ax[b] - xa) <= D) {
                    if (EMIT) pb[o] = b;
                    ++o;
                }
            }
        }
    if (!EMIT) cnt[a] = o;
}

// Per row entry j (row a, neighbour b = pb[j]): its row, whether it holds the
// pair's block (b >= a: the upper entries), and for an upper entry the entry
// of row b that lists a (rev; itself on the diagonal).  The overlap test is
// symmetric, so that entry exists; *bad is set if one is missing.
static __global__ void k_pair_rev(int na, const int* __restrict__ rowptr, const int* __restrict__ pb,
                                  int* __restrict__ rowof, uint8_t* __restrict__ upper, int* __restrict__ rev,
                                  int* __restrict__ bad) {
    const int a = blockIdx.x * BLOCK + threadIdx.x;
    if (a >= na) return;
    for (int j = rowptr[a]; j < rowptr[a + 1]; ++j) {
        const int b = pb[j];
        rowof[j] = a;
        upper[j] = b >= a;
        if (b == a) rev[j] = j;
        if (b >= a) continue;
        int t = -1;
        for (int i = rowptr[b]; i < rowptr[b + 1]; ++i)
            if (pb[i] == a) {
                t = i;
                break;
            }
        if (t < 0) *bad = 1;
        else rev[t] = j;   // row b's entry for a is the upper one
    }
}

// The record of upper entry u (slot j = ulist[u]): (j, its row, its
// neighbour, its twin entry); a row's own entry (neighbour = row) has no twin
// and carries the row's length instead.
static __global__ void k_upper_rec(int n_upper, const int* __restrict__ ulist, const int* __restrict__ rowof,
                                   const int* __restrict__ pb, const int* __restrict__ rev,
                                   const int* __restrict__ rowptr, int4* __restrict__ urec) {
    const int u = blockIdx.x * BLOCK + threadIdx.x;
    if (u >= n_upper) return;
    const int j = ulist[u], a = rowof[j], b = pb[j];
    urec[u] = make_int4(j, a, b, b == a ? rowptr[a + 1] - rowptr[a] : rev[j]);
}

// Entry words (neighbour, code) of the rows live at the last compaction,
// packed row after row (entry j of such a row at cpos[j], row a from cst[a]),
// rewritten by every H assembly: the neighbour is -1 where either star is
// frozen (the neighbour's vectors and the block are not read; a frozen row's
// entries are all -1).  The code is the slot j (>= 0) of the pair's block,
// -(j + 1) where the row reads the block of the pair's upper entry j
// transposed, or ENT_HEAD + the row's length for the row's own entry, the
// first of the row, whose block is the star's diagonal block Hd[a].  Slots
// stay below 2^30, so the three ranges do not meet.
constexpr int ENT_HEAD = INT_MIN;
constexpr int ENT_HEAD_MAX = -(1 << 30) - 1;   // codes <= this are row heads

// H_ab (P x P, row-major: row k is star a's parameter k) of every upper entry
// (its record: slot j, row a, neighbour b >= a, the twin entry of row b) whose
// two stars are live: GH lanes per entry
// striding over the pixels the two stamps share, the P^2 sums as group
// reductions in float64, stored in slot j.  Every upper entry also writes the
// (neighbour, slot) words of both rows' entries for this outer iteration:
// the entry words above.
template <typename T, int GH, int P>
__global__ __launch_bounds__(BLOCK) void k_hbuild(const int4* __restrict__ urec, int n_upper,
                                                  const uint8_t* __restrict__ live, const int* __restrict__ ay,
                                                  const int* __restrict__ ax, int R, int S,
                                                  const T* __restrict__ V, double* __restrict__ H,
                                                  double* __restrict__ Hd, const int* __restrict__ cpos,
                                                  int2* __restrict__ ent) {
    constexpr int GPB = BLOCK / GH, PP = P * P;
    const int sub = threadIdx.x % GH;
    const int S2 = S * S;
    const vec4_t<T>* V4 = reinterpret_cast<const vec4_t<T>*>(V);
    for (int u = xcd_remap(blockIdx.x, gridDim.x) * GPB + threadIdx.x / GH; u < n_upper; u += gridDim.x * GPB) {
        const int4 ur = urec[u];   // (slot j, row a, neighbour b, twin entry)
        const int j = ur.x, a = ur.y, b = ur.z;
        const bool la = live[a], lb = live[b];
        const int dy0 = ay[b] - ay[a], dx0 = ax[b] - ax[a];
        if (sub == 0) {
            const int cj = cpos[j];
            if (b == a) {
                if (cj >= 0) ent[cj] = make_int2(la ? a : -1, ENT_HEAD + ur.w);
            } else {
                const int ct = cpos[ur.w];
                if (cj >= 0) ent[cj] = make_int2(la && lb ? b : -1, j);
                if (ct >= 0) ent[ct] = make_int2(la && lb ? a : -1, -(j + 1));
            }
        }
        if (!(la && lb)) continue;   // uniform within the group
        double h[PP];
#pragma unroll
        for (int e = 0; e < PP; ++e) h[e] = 0;
        const int ylo = max(-R, dy0 - R), yhi = min(R, dy0 + R);
        const int xlo = max(-R, dx0 - R), xhi = min(R, dx0 + R);
        const int hy = yhi - ylo + 1, nov = hy * (xhi - xlo + 1);
        // pixel i of the overlap is (ylo + i % hy, xlo + i / hy); step GH
        const int qs = GH / hy, rs = GH % hy;
        int oy = sub % hy, ox = sub / hy;
        const vec4_t<T>* Va = V4 + (size_t)a * S2 + R * S + R;
        const vec4_t<T>* Vb = V4 + (size_t)b * S2 + (R - dx0) * S + (R - dy0);
        for (int i = sub; i < nov; i += GH) {
            const int dy = ylo + oy, dx = xlo + ox;
            const vec4_t<T> va = Va[dx * S + dy];
            const vec4_t<T> vb = Vb[dx * S + dy];
            const double A[3] = {(double)va.x, (double)va.y, (double)va.z};
            const double B[3] = {(double)vb.x, (double)vb.y, (double)vb.z};
#pragma unroll
            for (int k = 0; k < P; ++k)
#pragma unroll
                for (int l = 0; l < P; ++l) h[P * k + l] += A[k] * B[l];
            oy += rs;
            ox += qs;
            if (oy >= hy) {
                oy -= hy;
                ++ox;
            }
        }
#pragma unroll
        for (int e = 0; e < PP; ++e) h[e] = group_sum<GH>(h[e]);
        // lane e < P^2 writes entry e of the block (GH < P^2: lanes write
        // several); a star's diagonal block goes to Hd
        double* hout = b == a ? Hd + (size_t)a * PP : H + (size_t)j * PP;
        for (int e0 = sub; e0 < PP; e0 += GH) {
            double v = 0;
#pragma unroll
            for (int e = 0; e < PP; ++e)
                if (e == e0) v = h[e];
            hout[e0] = v;
        }
    }
}

// Block sums of N values in a fixed order (wave shuffles, then thread k sums
// the waves of value k); value k of this block goes to part[k * gridDim.x +
// blockIdx.x] with a plain store, read by the one-block finish launch after the
// kernel boundary.
template <int N>
__device__ __forceinline__ void block_store_n(double (&v)[N], double* __restrict__ part, double (*sh)[NWAVE]) {
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] = wave_sum(v[k]);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < N; ++k) sh[k][wv] = v[k];
    __syncthreads();
    if (threadIdx.x < N) {
        double r = 0;
        for (int i = 0; i < NWAVE; ++i) r += sh[threadIdx.x][i];
        part[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = r;
    }
}

// Sum over the 4 lanes of an aligned quad (DPP lane swaps; every lane gets
// the same value).
template <int CTRL>
__device__ __forceinline__ double dpp_swap(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b & 0xffffffffll), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ __forceinline__ double quad_sum(double v) {
    v += dpp_swap<0x4E>(v);   // quad_perm [2, 3, 0, 1]: lane ^ 2
    v += dpp_swap<0xB1>(v);   // quad_perm [1, 0, 3, 2]: lane ^ 1
    return v;
}

// y += H_ab v_b over the live neighbours b of row a, with v_b = c1 s1[b] +
// c2 s2[b] (s2 read only when TWO): the 4 lanes of a quad take the row's
// entries j = jb, jb + 4, ... below j1, and the quad sums y.  The row's own
// entry (b = a) also yields s1[a] and v_a (sd, vd), so the row needs no
// separate loads of them; every lane gets all three.
template <int P, bool TWO>
__device__ __forceinline__ void row_apply(int a, int jb, int j1, const int2* __restrict__ ent,
                                          const double* __restrict__ H, const double* __restrict__ Hd,
                                          const double* __restrict__ s1, int st1,
                                          double c1, const double* __restrict__ s2, double c2, double (&y)[P],
                                          double (&sd)[P], double (&vd)[P]) {
    constexpr int PP = P * P;
    const unsigned long long* ew = reinterpret_cast<const unsigned long long*>(ent);
#pragma unroll
    for (int k = 0; k < P; ++k) y[k] = sd[k] = vd[k] = 0;
    for (int j = jb; j < j1; j += 4) {
        // one 8-byte load of (neighbour, code); a frozen neighbour (-1) reads
        // star 0 and its block slot (in bounds) and contributes nothing
        const unsigned long long e = ew[j];
        const int bn = (int)(unsigned)(e & 0xffffffffu), cd = (int)(unsigned)(e >> 32);
        const bool on = bn >= 0, dg = bn == a;
        const int b = on ? bn : 0;
        const bool tr = cd < 0 && cd > ENT_HEAD_MAX;
        const bool hd = cd <= ENT_HEAD_MAX;
        double v[P];
#pragma unroll
        for (int l = 0; l < P; ++l) {
            const double sv = s1[(size_t)b * st1 + l];
            v[l] = c1 * sv;
            if (TWO) v[l] += c2 * s2[(size_t)b * P + l];
            sd[l] = dg ? sv : sd[l];
            vd[l] = dg ? v[l] : vd[l];
        }
        const double* h = hd ? Hd + (size_t)b * PP : H + (size_t)(tr ? -cd - 1 : cd) * PP;
        double hv[PP];
#pragma unroll
        for (int q = 0; q < PP; ++q) hv[q] = h[q];
#pragma unroll
        for (int k = 0; k < P; ++k) {
            double t = 0;
#pragma unroll
            for (int l = 0; l < P; ++l) t += (tr ? hv[P * l + k] : hv[P * k + l]) * v[l];
            y[k] += on ? t : 0.0;
        }
    }
#pragma unroll
    for (int k = 0; k < P; ++k) {
        y[k] = quad_sum(y[k]);
        sd[k] = quad_sum(sd[k]);   // one lane holds it, the others 0: exact
        vd[k] = quad_sum(vd[k]);
    }
}

// y[k] for a lane-dependent k as explicit selects over registers (an indexed
// read would move the array to LDS).
template <int P>
__device__ __forceinline__ double pick(const double (&y)[P], int k) {
    const double y0 = y[0], y1 = P > 1 ? y[1] : 0.0, y2 = P > 2 ? y[2] : 0.0;
    return k == 0 ? y0 : (k == 1 ? y1 : y2);
}

// Lanes per row of CG's H p (a quad: row_apply's DPP sums).
constexpr int GKL = 4;

// The scalar state of a fit staged in LDS (as 8-byte words, threads < NW)
// and written back.
constexpr int SCAL_WORDS = sizeof(Scal) / 8;
static_assert(sizeof(Scal) % 8 == 0, "Scal staged as 8-byte words");
__device__ __forceinline__ void scal_stage(Scal& Ls, const Scal* Sc, int tid) {
    if (tid < SCAL_WORDS)
        reinterpret_cast<unsigned long long*>(&Ls)[tid] = reinterpret_cast<const unsigned long long*>(Sc)[tid];
}
__device__ __forceinline__ void scal_store(Scal* Sc, const Scal& Ls, int tid) {
    if (tid < SCAL_WORDS)
        reinterpret_cast<unsigned long long*>(Sc)[tid] = reinterpret_cast<const unsigned long long*>(&Ls)[tid];
}

// The scalars of Golub-Kahan step it from its partials (nb blocks, NP each):
// their block-order sum, beta_{k+1} and alpha_{k+1} by the expansion above,
// then the LSQR or LSMR recurrences of step it, applied to the staged state
// Ls (left unchanged once the solve is done).  Threads t < FINGK_BLOCK sum
// the blocks t, t + FINGK_BLOCK, ..., so every caller sums in one order;
// threads past them only pass the barriers.  Ends with a barrier.
constexpr int FINGK_BLOCK = 512;
static_assert(SCAL_WORDS <= FINGK_BLOCK, "the scalar state is staged by one word per thread");

// Thread 0's part of gk_scalars (sh: the NP x 8 wave sums).  No fused
// multiply-adds here or in the recurrences it calls, so the reduction kernel
// and a folded step round alike wherever the code is inlined.
template <int SOLVER>
__device__ __forceinline__ void gk_advance(Scal* Ls, const double* sh, int it) {
#pragma clang fp contract(off)
    constexpr bool LSMR = SOLVER == CP_SOLVER_LSMR;
    constexpr int NP = LSMR ? 10 : 5;
    double t[NP];
    for (int e = 0; e < NP; ++e) {
        t[e] = 0;
        for (int i = 0; i < FINGK_BLOCK / 64; ++i) t[e] += sh[e * (FINGK_BLOCK / 64) + i];
    }
    const double al = Ls->alpha;
    const double beta = sqrt(fmax(0.0, t[0] - 2 * al * t[1] + al * al));
    Ls->beta_next = beta;
    const double gv = t[0] - al * t[1];
    const double tt = beta > 0 ? fmax(0.0, t[2] / (beta * beta) - 2 * gv + beta * beta * t[3]) : 0.0;
    if (LSMR) {
        // |x|^2, x.h, x.hbar, |h|^2, h.hbar, |hbar|^2
        const double q[6] = {t[5], t[6], t[7], t[4], t[8], t[9]};
        lsmr_advance(Ls, tt, q, it);
    } else {
        lsqr_advance(Ls, tt, t[4], it);
    }
}

template <int SOLVER>
__device__ __forceinline__ void gk_scalars(Scal& Ls, const double* __restrict__ part, int nb, int it, int tid) {
    constexpr bool LSMR = SOLVER == CP_SOLVER_LSMR;
    constexpr int NP = LSMR ? 10 : 5;
    __shared__ double sh[NP][FINGK_BLOCK / 64];
    double a[NP];
#pragma unroll
    for (int e = 0; e < NP; ++e) a[e] = 0;
    if (tid < FINGK_BLOCK)
        for (int i = tid; i < nb; i += FINGK_BLOCK)
#pragma unroll
            for (int e = 0; e < NP; ++e) a[e] += part[(size_t)e * nb + i];
#pragma unroll
    for (int e = 0; e < NP; ++e) a[e] = wave_sum(a[e]);
    if ((tid & 63) == 0 && tid < FINGK_BLOCK)
#pragma unroll
        for (int e = 0; e < NP; ++e) sh[e][tid >> 6] = a[e];
    __syncthreads();
    if (tid == 0 && !Ls.done) gk_advance<SOLVER>(&Ls, &sh[0][0], it);
    __syncthreads();
}

// One Golub-Kahan step it of LSQR / LSMR, one lane per row entry.
//
// Rows are consecutive entry ranges headed by the row's own entry (b = a,
// its code carries the row's length).
// Block k takes the GKB entries from k * hspan on and sums the rows whose
// head lies in its first hspan entries; the remaining GKB - hspan lanes only
// gather for the rows that run past that span (a row longer than the margin
// and headed near the span's end has its last entries summed by its head
// from global memory).  Each lane forms its neighbour's v_it where it
// gathers it (v_it = c1 s1[b] + c2 vp[b]), multiplies by the entry's block
// (read transposed where the slot is negative) and leaves H_ab v_b in LDS;
// the head of a live row sums its row, forms g_it = H v_it - alpha J^T u_it
// and v_it, and applies the pending x / w (LSMR: hbar) update of step it - 1.
// The first step forms v_1 = -b_scaled / (beta_1 alpha_1) and
// J^T u_1 = -b_scaled / beta_1.  Frozen rows have every entry word -1, so
// their lanes gather nothing.
// Partials: v.y, v.J^T u, |g|^2, |v|^2, |w|^2 and, for LSMR, |x|^2, x.w,
// x.hbar, w.hbar, |hbar|^2 of the updated vectors.
// Load order matters: vector loads return in issue order, so the entry word
// goes first, then the gather and, for a head, its row's x / w / hbar.
// Two 1024-thread blocks per CU (64 registers).
//
// FOLD (grids of at most one block per CU, one 1024-thread block per CU at
// 128 registers): the step also does the previous step's reduction, so a
// solve is one launch per step plus one reduction at its end.  Every block
// reduces the previous step's partials (part_prev, in the one order of
// gk_scalars) from the state Sc before that reduction into LDS and runs on the
// result; block 0 leaves the result in Sc_next for the next step.  The state
// and the partials alternate between two buffers, so no block reads what
// another block of the same launch writes.  The first step reduces nothing
// and passes Sc on.
constexpr int GKB = 1024;

template <int P, int SOLVER, bool FOLD>
__global__ __launch_bounds__(GKB) __attribute__((amdgpu_waves_per_eu(FOLD ? 4 : 8, FOLD ? 4 : 8))) void k_gk(
    int hspan, const int2* __restrict__ ent, const double* __restrict__ H, const double* __restrict__ Hd,
    const double* __restrict__ b_scaled, const double* __restrict__ gp, const double* __restrict__ vp,
    double* __restrict__ gn, double* __restrict__ vn, double* __restrict__ x, double* __restrict__ wv,
    double* __restrict__ hb, int first, double* __restrict__ part, const Scal* __restrict__ Sc,
    const double* __restrict__ part_prev, Scal* __restrict__ Sc_next, int it) {
    constexpr bool LSMR = SOLVER == CP_SOLVER_LSMR;
    constexpr int NP = LSMR ? 10 : 5, PP = P * P;
    __shared__ double sh[NP][GKB / 64];
    __shared__ double ty[P][GKB];   // H_ab v_b of every lane's entry
    const int l = threadIdx.x;
    const Scal* S = Sc;
    // the packed entries of the rows live at the last compaction (ne, fixed
    // within a solve) take the first nbk blocks (dispatched over every XCD,
    // each XCD's blocks renumbered to one contiguous range); the other blocks
    // and a finished solve do nothing.  FOLD: they find out after the fold,
    // whose reduction runs while the block's gathers are in flight.
    if constexpr (!FOLD)
        if (Sc->done) return;   // uniform (the one-block reduction then does nothing)
    const int ne = Sc->n_gk_entries;
    const int nbk = (int)(((long long)ne + hspan - 1) / hspan);
    const bool lblk = (int)blockIdx.x < nbk;   // uniform
    if constexpr (!FOLD) {
        if (!lblk) {
            if (l < NP) part[(size_t)l * gridDim.x + blockIdx.x] = 0.0;
            return;
        }
    }
    const int blk = lblk ? xcd_remap(blockIdx.x, nbk) : 0;
    const long long j = (long long)blk * hspan + l;
    const bool inr = lblk && j < ne;
    const unsigned long long* ewp = reinterpret_cast<const unsigned long long*>(ent);
    const size_t jc = (size_t)(inr ? j : (ne > 0 ? ne - 1 : 0));
    // the entry word (at a clamped, in-bounds index)
    const unsigned long long ew = ewp[jc];
    const int bn = (int)(unsigned)(ew & 0xffffffffu), cd = (int)(unsigned)(ew >> 32);
    const bool on = inr && bn >= 0;
    const bool hd = cd <= ENT_HEAD_MAX;
    const int L = (on && hd && l < hspan) ? cd - ENT_HEAD : 0;
    const bool head = L > 0;   // a live row's own entry (b = a) in this block's span
    const int b = on ? bn : 0;
    const bool tr = cd < 0 && !hd;
    const double* hsrc = hd ? Hd + (size_t)b * PP : H + (on ? (size_t)(tr ? -cd - 1 : cd) : 0) * PP;
    const double* s1 = first ? b_scaled : gp;
    const int st1 = first ? VP : P;
    double s1v[P], s2v[P], hv[PP];
#pragma unroll
    for (int k = 0; k < P; ++k) s1v[k] = s1[(size_t)b * st1 + k];
#pragma unroll
    for (int k = 0; k < P; ++k) s2v[k] = first ? 0.0 : vp[(size_t)b * P + k];
#pragma unroll
    for (int q = 0; q < PP; ++q) hv[q] = hsrc[q];
    if constexpr (FOLD) {
        __shared__ Scal Ls;
        scal_stage(Ls, Sc, l);
        if (first) __syncthreads();
        else gk_scalars<SOLVER>(Ls, part_prev, gridDim.x, it - 1, l);
        if (blockIdx.x == 0) scal_store(Sc_next, Ls, l);
        S = &Ls;
        if (Ls.done) return;   // uniform
        if (!lblk) {
            if (l < NP) part[(size_t)l * gridDim.x + blockIdx.x] = 0.0;
            return;
        }
    }
    const size_t rb = (size_t)b * P;
    const double alpha = S->alpha;
    const double ia = alpha != 0 ? 1.0 / alpha : 1.0;
    const double beta = first ? 0.0 : S->beta_cur;
    const double ib = beta != 0 ? 1.0 / beta : 1.0;
    const double ib1 = first ? 1.0 / S->beta1 : 0.0;
    // v_it = c1 s1 + c2 vp, J^T u_it = d1 s1
    const double c1 = first ? -ib1 * ia : ib * ia, c2 = first ? 0.0 : -beta * ia, d1 = first ? -ib1 : ib;
    const double sig = first ? 0.0 : S->sig, tau = first ? 0.0 : S->tau, del = first ? 0.0 : S->del;
    double v[P], y[P];
#pragma unroll
    for (int k = 0; k < P; ++k) v[k] = c1 * s1v[k] + c2 * s2v[k];
#pragma unroll
    for (int k = 0; k < P; ++k) {
        double s = 0;
#pragma unroll
        for (int q = 0; q < P; ++q) s += (tr ? hv[P * q + k] : hv[P * k + q]) * v[q];
        y[k] = on ? s : 0.0;
        ty[k][l] = y[k];
    }
    __syncthreads();
    double acc[NP];
#pragma unroll
    for (int e = 0; e < NP; ++e) acc[e] = 0;
    if (head) {
        // (no fused multiply-adds here: the unfolded and the folded kernel then
        // round the row's updates alike, whatever their register budgets make
        // of the code)
#pragma clang fp contract(off)
        // the row's x / w / hbar, loaded while the row is summed (loaded with
        // the gather they would hold registers the gather needs: the step runs
        // at 64 registers so two blocks share a CU)
        double xo[P], wo[P], ho[P];
#pragma unroll
        for (int k = 0; k < P; ++k) {
            xo[k] = !first ? x[rb + k] : 0.0;
            wo[k] = !first ? wv[rb + k] : 0.0;
            ho[k] = (LSMR && !first) ? hb[rb + k] : 0.0;
        }
        // the row's other entries in order: from LDS within the block, past
        // it from global memory
        const int lend = min(l + L, GKB);
        for (int i = l + 1; i < lend; ++i)
#pragma unroll
            for (int k = 0; k < P; ++k) y[k] += ty[k][i];
        for (long long jj = j + (lend - l); jj < j + L; ++jj) {
            const unsigned long long e2 = ewp[jj];
            const int b2n = (int)(unsigned)(e2 & 0xffffffffu), s2l = (int)(unsigned)(e2 >> 32);
            if (b2n < 0) continue;   // (a row's only head is its first entry)
            const bool t2 = s2l < 0;
            const double* h2 = H + (size_t)(t2 ? -s2l - 1 : s2l) * PP;
            double v2[P];
#pragma unroll
            for (int k = 0; k < P; ++k)
                v2[k] = c1 * s1[(size_t)b2n * st1 + k] + (first ? 0.0 : c2 * vp[(size_t)b2n * P + k]);
#pragma unroll
            for (int k = 0; k < P; ++k)
#pragma unroll
                for (int q = 0; q < P; ++q) y[k] += (t2 ? h2[P * q + k] : h2[P * k + q]) * v2[q];
        }
#pragma unroll
        for (int k = 0; k < P; ++k) {
            const double yk = y[k], vk = v[k];   // the head's own v_a
            const double jk = d1 * s1v[k];       // and J^T u_a
            const double gk = yk - alpha * jk;
            gn[rb + k] = gk;
            vn[rb + k] = vk;
            acc[0] += vk * yk;
            acc[1] += vk * jk;
            acc[2] += gk * gk;
            acc[3] += vk * vk;
            const double hn = wo[k] - del * ho[k];
            const double xn = xo[k] + sig * hn;
            const double wn = vk - tau * wo[k];
            x[rb + k] = xn;
            wv[rb + k] = wn;
            acc[4] += wn * wn;
            if (LSMR) {
                hb[rb + k] = hn;
                acc[5] += xn * xn;
                acc[6] += xn * wn;
                acc[7] += xn * hn;
                acc[8] += wn * hn;
                acc[9] += hn * hn;
            }
        }
    }
#pragma unroll
    for (int e = 0; e < NP; ++e) acc[e] = wave_sum(acc[e]);
    if ((l & 63) == 0)
#pragma unroll
        for (int e = 0; e < NP; ++e) sh[e][l >> 6] = acc[e];
    __syncthreads();
    if (l < NP) {
        double r = 0;
        for (int w = 0; w < GKB / 64; ++w) r += sh[l][w];
        part[(size_t)l * gridDim.x + blk] = r;
    }
}

// One-block reduction of Golub-Kahan step it (after k_gk, or after the last
// folded step): the state in Sc_in advanced by the step's partials, written to
// Sc_out.  The scalar state is staged in LDS while the partials load, so the
// serial recurrences read and write LDS, not global memory.
// (512 threads: the LSMR recurrences need more than the 128 registers a
// 1024-thread block leaves each lane)
template <int SOLVER>
static __global__ __launch_bounds__(FINGK_BLOCK) void k_finish_gk(const Scal* Sc_in, Scal* Sc_out,
                                                                  const double* __restrict__ part, int nb, int it) {
    __shared__ Scal Ls;
    scal_stage(Ls, Sc_in, threadIdx.x);
    gk_scalars<SOLVER>(Ls, part, nb, it, threadIdx.x);
    scal_store(Sc_out, Ls, threadIdx.x);
}

// The end of a folded solve: every block reduces the last step's partials
// (nb blocks, step it) from the state Sc_in before that reduction, block 0
// writes the result to Sc_out, and each thread takes one star's final x
// update, capped step and candidate theta (step_star, as k_step after
// k_finish_gk).
template <int SOLVER>
static __global__ __launch_bounds__(FINGK_BLOCK) void k_step_fold(
    int na, int p, int k_y, int k_x, double max_step, const double* __restrict__ wv, const double* __restrict__ hb,
    double* __restrict__ x, const double* __restrict__ colnorm, const double* __restrict__ theta,
    double* __restrict__ dsc, double* __restrict__ theta_c, const uint8_t* __restrict__ live,
    const Scal* __restrict__ Sc_in, Scal* __restrict__ Sc_out, const double* __restrict__ part, int nb, int it) {
    __shared__ Scal Ls;
    scal_stage(Ls, Sc_in, threadIdx.x);
    gk_scalars<SOLVER>(Ls, part, nb, it, threadIdx.x);
    if (blockIdx.x == 0) scal_store(Sc_out, Ls, threadIdx.x);
    const int a = blockIdx.x * FINGK_BLOCK + threadIdx.x;
    if (a < na) step_star(a, p, k_y, k_x, max_step, wv, hb, x, colnorm, theta, dsc, theta_c, live, &Ls);
}

// Packing the entries of the live rows (at every compaction of the live
// list): the row lengths of live stars (0 for frozen ones, and for a = na),
// whose exclusive scan gives each live row's start cst[a]; then every entry
// j of row a gets cpos[j] = cst[a] + (j - rowptr[a]), or -1 for a frozen
// row, and Sc->n_gk_entries the total.
static __global__ void k_live_rowlen(int na, const uint8_t* __restrict__ live, const int* __restrict__ rowptr,
                                     int* __restrict__ rlen) {
    const int a = blockIdx.x * BLOCK + threadIdx.x;
    if (a <= na) rlen[a] = (a < na && live[a]) ? rowptr[a + 1] - rowptr[a] : 0;
}
static __global__ void k_live_cpos(int na, const uint8_t* __restrict__ live, const int* __restrict__ rowptr,
                                   const int* __restrict__ cst, int* __restrict__ cpos, Scal* Sc) {
    const int a = blockIdx.x * BLOCK + threadIdx.x;
    if (a == 0) Sc->n_gk_entries = cst[na];
    if (a >= na) return;
    const int j0 = rowptr[a], j1 = rowptr[a + 1], c0 = live[a] ? cst[a] : -1;
    for (int j = j0; j < j1; ++j) cpos[j] = c0 >= 0 ? c0 + (j - j0) : -1;
}

// CG on the damped normal equations (J'J + lam^2 I) x = -b_scaled, the
// direction half: q = H p, a quad of lanes per star row (a quad for every
// star, lane k < p owning component k) with p = s + beta p_old formed where
// it is gathered (the first step: p_0 = s_0 = -b_scaled); each live row
// writes its own p to pn and q.  Partials p.q and |p|^2.
template <int P>
__global__ __launch_bounds__(BLOCK) void k_cg_hv(int na, const uint8_t* __restrict__ live,
                                                 const int* __restrict__ rowptr, const int* __restrict__ cst,
                                                 const int2* __restrict__ ent, const double* __restrict__ H,
                                                 const double* __restrict__ Hd, const double* __restrict__ b_scaled,
                                                 const double* __restrict__ s, const double* __restrict__ po,
                                                 double* __restrict__ pn, double* __restrict__ q, int first,
                                                 double* __restrict__ part, Scal* Sc) {
    constexpr int RPB = BLOCK / GKL;
    __shared__ double sred[2][NWAVE];
    const int sub = threadIdx.x % GKL;
    const int a = xcd_remap(blockIdx.x, gridDim.x) * RPB + threadIdx.x / GKL;
    const double* s1 = first ? b_scaled : s;
    const int st1 = first ? VP : P;
    // one combined row test over loads issued together, as in k_gk
    const int ac = a < na ? a : na - 1;
    const size_t ck = (size_t)ac * P + (sub < P ? sub : P - 1);
    // a live row is one of the rows packed at the last compaction, from cst[a]
    const int j0 = cst[ac], j1 = j0 + rowptr[ac + 1] - rowptr[ac];
    const bool row = (a < na) & (live[ac] != 0) & (j1 > j0) & (Sc->done == 0);
    const double cb = first ? 0.0 : Sc->cg_beta;
    const double c1 = first ? -1.0 : 1.0;
    double acc[2] = {0, 0};
    if (row) {
        double y[P], sd[P], pd[P];
        if (first) row_apply<P, false>(a, j0 + sub, j1, ent, H, Hd, s1, st1, c1, po, cb, y, sd, pd);
        else row_apply<P, true>(a, j0 + sub, j1, ent, H, Hd, s1, st1, c1, po, cb, y, sd, pd);
        if (sub < P) {
            const double yk = pick<P>(y, sub), pk = pick<P>(pd, sub);
            pn[ck] = pk;
            q[ck] = yk;
            acc[0] += pk * yk;
            acc[1] += pk * pk;
        }
    }
    block_store_n<2>(acc, part, sred);
}

// CG, the update half (one thread per star of lst): x <- x + a p,
// s <- s - a (q + lam^2 p) (s_0 = -b_scaled).  Partials |s|^2 and s.p.
template <int P>
__global__ __launch_bounds__(BLOCK) void k_cg_update(const int* __restrict__ lst, const uint8_t* __restrict__ live,
                                                     const double* __restrict__ b_scaled,
                                                     const double* __restrict__ pc, const double* __restrict__ q,
                                                     double* __restrict__ s, double* __restrict__ x, int first,
                                                     double* __restrict__ part, Scal* Sc) {
    if (Sc->done) return;
    __shared__ double sred[2][NWAVE];
    const double ca = Sc->cg_alpha, lam2 = Sc->lam * Sc->lam;
    double acc[2] = {0, 0};
    const int nl = Sc->n_jtu_stars;
    for (int li = xcd_remap(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x; li < nl; li += gridDim.x * BLOCK) {
        const int a = lst[li];
        const size_t base = (size_t)a * P;
        if (!live[a]) {
            if (first)
#pragma unroll
                for (int k = 0; k < P; ++k) x[base + k] = 0;
            continue;
        }
#pragma unroll
        for (int k = 0; k < P; ++k) {
            const double pk = pc[base + k];
            x[base + k] = (first ? 0.0 : x[base + k]) + ca * pk;
            const double sk = first ? -b_scaled[(size_t)a * VP + k] : s[base + k];
            const double sn = sk - ca * (q[base + k] + lam2 * pk);
            s[base + k] = sn;
            acc[0] += sn * sn;
            acc[1] += sn * pk;
        }
    }
    block_store_n<2>(acc, part, sred);
}

}  // namespace cp
