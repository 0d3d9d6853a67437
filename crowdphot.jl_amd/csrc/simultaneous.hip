// simultaneous.hip -- host orchestration of fit_all_stars_simultaneous on
// gfx950 (src/photometry/psf_photometry_simultaneous.jl:953-1431).
//
// Control flow mirrors the reference outer loop: per outer iteration one
// fused render+residual pass, one fill+J^T r pass, then damping trials of
// (inner LSQR solve, capped step, candidate cost) until a trial lowers the
// cost; accepted steps run the per-star freeze tests and subtract newly
// frozen stars from the working image.  The host reads back a few scalars per
// trial (cost, cost_cand, counts) to take the same accept / converge
// decisions; everything else stays on the device.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <type_traits>
#include <cmath>
#include <cstring>
#include <vector>

#include "aperture_host.h"
#include "diag_host.h"
#include "normal_op.h"
#include "sim_kernels.h"
#include "psf_host.h"

namespace cp {

// ---------------------------------------------------------------------------
// hipcub helpers
// ---------------------------------------------------------------------------
struct Cub {
    cp_ctx* ctx;
    DevBuf tmp;
    void ensure(size_t b) {
        if (tmp.bytes() < b) tmp.alloc(ctx, std::max<size_t>(b, 1 << 20));
    }
};

// Selected (ok) catalogue ids, stably sorted by Morton key: the active list.
int build_active(cp_ctx* ctx, Cub& cub, const int* d_ay, const int* d_ax, const uint8_t* d_ok, int n,
                 int coord_bits, DevBuf& act_out) {
    hipStream_t st = ctx->stream;
    DevBuf sel(ctx, (size_t)std::max(n, 1) * 4), nsel(ctx, 8);
    hipcub::CountingInputIterator<int> it(0);
    size_t b = 0;
    CP_HIP(hipcub::DeviceSelect::Flagged(nullptr, b, it, d_ok, sel.as<int>(), nsel.as<int>(), n, st));
    cub.ensure(b);
    b = cub.tmp.bytes();
    CP_HIP(hipcub::DeviceSelect::Flagged(cub.tmp.get(), b, it, d_ok, sel.as<int>(), nsel.as<int>(), n, st));
    int na = 0;
    CP_HIP(hipMemcpyAsync(&na, nsel.get(), 4, hipMemcpyDeviceToHost, st));
    CP_HIP(hipStreamSynchronize(st));
    act_out.alloc(ctx, (size_t)std::max(na, 1) * 4);
    if (na == 0) return 0;
    DevBuf keys(ctx, (size_t)na * 8), keys2(ctx, (size_t)na * 8);
    launch(k_morton_keys, dim3(cdiv(na, BLOCK)), dim3(BLOCK), 0, st, (const int*)sel.as<int>(), na, d_ay, d_ax,
           keys.as<unsigned long long>());
    int end_bit = std::min(64, 2 * coord_bits);
    b = 0;
    CP_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, b, keys.as<unsigned long long>(),
                                              keys2.as<unsigned long long>(), sel.as<int>(), act_out.as<int>(),
                                              na, 0, end_bit, st));
    cub.ensure(b);
    b = cub.tmp.bytes();
    CP_HIP(hipcub::DeviceRadixSort::SortPairs(cub.tmp.get(), b, keys.as<unsigned long long>(),
                                              keys2.as<unsigned long long>(), sel.as<int>(), act_out.as<int>(),
                                              na, 0, end_bit, st));
    return na;
}

struct TileLists {
    DevBuf start, end, list;
    int npairs = 0;
};

// Stars whose box (half-width hw[a] or h) overlaps each tile, in star order.
void build_tiles(cp_ctx* ctx, Cub& cub, int na, const int* d_ay, const int* d_ax, const int* d_hw, int h,
                 const TileGeo& g, TileLists& L) {
    hipStream_t st = ctx->stream;
    const int ntiles = g.nty * g.ntx;
    DevBuf cnt(ctx, (size_t)na * 4), off(ctx, (size_t)na * 4), tot(ctx, 8);
    launch(k_tile_count, dim3(cdiv(na, BLOCK)), dim3(BLOCK), 0, st, na, d_ay, d_ax, d_hw, h, g, cnt.as<int>());
    size_t b = 0;
    CP_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, b, cnt.as<int>(), off.as<int>(), na, st));
    cub.ensure(b);
    b = cub.tmp.bytes();
    CP_HIP(hipcub::DeviceScan::ExclusiveSum(cub.tmp.get(), b, cnt.as<int>(), off.as<int>(), na, st));
    int last_off = 0, last_cnt = 0;
    CP_HIP(hipMemcpyAsync(&last_off, off.as<int>() + na - 1, 4, hipMemcpyDeviceToHost, st));
    CP_HIP(hipMemcpyAsync(&last_cnt, cnt.as<int>() + na - 1, 4, hipMemcpyDeviceToHost, st));
    CP_HIP(hipStreamSynchronize(st));
    const int np = last_off + last_cnt;
    L.npairs = np;
    L.start.alloc(ctx, (size_t)ntiles * 4);
    L.end.alloc(ctx, (size_t)ntiles * 4);
    CP_HIP(hipMemsetAsync(L.start.get(), 0, (size_t)ntiles * 4, st));
    CP_HIP(hipMemsetAsync(L.end.get(), 0, (size_t)ntiles * 4, st));
    L.list.alloc(ctx, (size_t)std::max(np, 1) * 4);
    if (np == 0) return;
    DevBuf k1(ctx, (size_t)np * 4), k2(ctx, (size_t)np * 4), v1(ctx, (size_t)np * 4);
    launch(k_tile_emit, dim3(cdiv(na, BLOCK)), dim3(BLOCK), 0, st, na, d_ay, d_ax, d_hw, h, g,
           (const int*)off.as<int>(), k1.as<unsigned int>(), v1.as<int>());
    int end_bit = 1;
    while ((1LL << end_bit) < ntiles) ++end_bit;
    b = 0;
    CP_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, b, k1.as<unsigned int>(), k2.as<unsigned int>(),
                                              v1.as<int>(), L.list.as<int>(), np, 0, end_bit, st));
    cub.ensure(b);
    b = cub.tmp.bytes();
    CP_HIP(hipcub::DeviceRadixSort::SortPairs(cub.tmp.get(), b, k1.as<unsigned int>(), k2.as<unsigned int>(),
                                              v1.as<int>(), L.list.as<int>(), np, 0, end_bit, st));
    launch(k_tile_ranges, dim3(cdiv(np, BLOCK)), dim3(BLOCK), 0, st, (const unsigned int*)k2.as<unsigned int>(),
           np, L.start.as<int>(), L.end.as<int>());
}

// Rows of the star-space normal matrix: for every active star (Morton order)
// the stars whose stamps overlap its own, from a binning into cells of side 2R.
struct PairRows {
    DevBuf rowptr, pb;
    DevBuf urec;   // per upper entry (b >= a): (slot, row, neighbour, twin entry of the other row)
    int64_t npairs = 0;
    int n_upper = 0;
    int max_row = 0;
};

static int bit_length(int64_t v);

void build_pairs(cp_ctx* ctx, Cub& cub, int na, const int* d_ay, const int* d_ax, int R, int ny, int nx,
                 PairRows& P) {
    hipStream_t st = ctx->stream;
    CellGeo g{std::max(2 * R, 1), 0, 0, R};
    g.ncy = (ny + 2 * R - 1) / g.CS + 1;
    g.ncx = (nx + 2 * R - 1) / g.CS + 1;
    const int64_t ncells = (int64_t)g.ncy * g.ncx;
    CP_ARG(ncells < (int64_t)INT_MAX, "image too large for one fit (star cells exceed 2^31)");
    DevBuf k1(ctx, (size_t)na * 4), k2(ctx, (size_t)na * 4), v1(ctx, (size_t)na * 4), clist(ctx, (size_t)na * 4),
        cstart(ctx, (size_t)ncells * 4), cend(ctx, (size_t)ncells * 4), cnt(ctx, (size_t)na * 4);
    P.rowptr.alloc(ctx, (size_t)(na + 1) * 4);
    launch(k_cell_keys, dim3(cdiv(na, BLOCK)), dim3(BLOCK), 0, st, na, d_ay, d_ax, g, k1.as<unsigned int>(),
           v1.as<int>());
    const int end_bit = std::min(32, bit_length(ncells));
    size_t b = 0;
    CP_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, b, k1.as<unsigned int>(), k2.as<unsigned int>(), v1.as<int>(),
                                              clist.as<int>(), na, 0, end_bit, st));
    cub.ensure(b);
    b = cub.tmp.bytes();
    CP_HIP(hipcub::DeviceRadixSort::SortPairs(cub.tmp.get(), b, k1.as<unsigned int>(), k2.as<unsigned int>(),
                                              v1.as<int>(), clist.as<int>(), na, 0, end_bit, st));
    CP_HIP(hipMemsetAsync(cstart.get(), 0, (size_t)ncells * 4, st));
    CP_HIP(hipMemsetAsync(cend.get(), 0, (size_t)ncells * 4, st));
    launch(k_tile_ranges, dim3(cdiv(na, BLOCK)), dim3(BLOCK), 0, st, (const unsigned int*)k2.as<unsigned int>(), na,
           cstart.as<int>(), cend.as<int>());
    launch(k_pairs<false>, dim3(cdiv(na, BLOCK)), dim3(BLOCK), 0, st, na, d_ay, d_ax, g,
           (const int*)cstart.as<int>(), (const int*)cend.as<int>(), (const int*)clist.as<int>(), cnt.as<int>(),
           (const int*)nullptr, (int*)nullptr);
    b = 0;
    CP_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, b, cnt.as<int>(), P.rowptr.as<int>(), na, st));
    cub.ensure(b);
    b = cub.tmp.bytes();
    CP_HIP(hipcub::DeviceScan::ExclusiveSum(cub.tmp.get(), b, cnt.as<int>(), P.rowptr.as<int>(), na, st));
    // the longest row (the Golub-Kahan step's blocks overlap by its length)
    DevBuf mrow(ctx, 4);
    b = 0;
    CP_HIP(hipcub::DeviceReduce::Max(nullptr, b, cnt.as<int>(), mrow.as<int>(), na, st));
    cub.ensure(b);
    b = cub.tmp.bytes();
    CP_HIP(hipcub::DeviceReduce::Max(cub.tmp.get(), b, cnt.as<int>(), mrow.as<int>(), na, st));
    int hv[3] = {0, 0, 0};
    CP_HIP(hipMemcpyAsync(&hv[0], P.rowptr.as<int>() + na - 1, 4, hipMemcpyDeviceToHost, st));
    CP_HIP(hipMemcpyAsync(&hv[1], cnt.as<int>() + na - 1, 4, hipMemcpyDeviceToHost, st));
    CP_HIP(hipMemcpyAsync(&hv[2], mrow.get(), 4, hipMemcpyDeviceToHost, st));
    CP_HIP(hipStreamSynchronize(st));
    const int64_t np64 = (int64_t)hv[0] + hv[1];
    CP_ARG(np64 < ((int64_t)1 << 30), "too many overlapping star pairs for one fit (>= 2^30)");
    P.max_row = hv[2];
    P.npairs = np64;
    const int np32 = (int)np64;
    CP_HIP(hipMemcpyAsync(P.rowptr.as<int>() + na, &np32, 4, hipMemcpyHostToDevice, st));
    P.pb.alloc(ctx, (size_t)std::max<int64_t>(P.npairs, 1) * 4);
    launch(k_pairs<true>, dim3(cdiv(na, BLOCK)), dim3(BLOCK), 0, st, na, d_ay, d_ax, g,
           (const int*)cstart.as<int>(), (const int*)cend.as<int>(), (const int*)clist.as<int>(), (int*)nullptr,
           (const int*)P.rowptr.as<int>(), P.pb.as<int>());
    // each pair's block is kept once, in its upper entry (b >= a)
    const size_t ne = (size_t)std::max<int64_t>(P.npairs, 1);
    DevBuf rowof(ctx, ne * 4), rev(ctx, ne * 4), ulist(ctx, ne * 4), upper(ctx, ne), nup(ctx, 4);
    CP_HIP(hipMemsetAsync(cnt.get(), 0, 4, st));
    CP_HIP(hipMemsetAsync(upper.get(), 0, ne, st));
    launch(k_pair_rev, dim3(cdiv(na, BLOCK)), dim3(BLOCK), 0, st, na, (const int*)P.rowptr.as<int>(),
           (const int*)P.pb.as<int>(), rowof.as<int>(), upper.as<uint8_t>(), rev.as<int>(), cnt.as<int>());
    {
        hipcub::CountingInputIterator<int> cit(0);
        size_t b3 = 0;
        CP_HIP(hipcub::DeviceSelect::Flagged(nullptr, b3, cit, (const uint8_t*)upper.as<uint8_t>(), ulist.as<int>(),
                                             nup.as<int>(), np32, st));
        cub.ensure(b3);
        b3 = cub.tmp.bytes();
        CP_HIP(hipcub::DeviceSelect::Flagged(cub.tmp.get(), b3, cit, (const uint8_t*)upper.as<uint8_t>(),
                                             ulist.as<int>(), nup.as<int>(), np32, st));
    }
    int bad = 0;
    CP_HIP(hipMemcpyAsync(&bad, cnt.get(), 4, hipMemcpyDeviceToHost, st));
    CP_HIP(hipMemcpyAsync(&P.n_upper, nup.get(), 4, hipMemcpyDeviceToHost, st));
    CP_HIP(hipStreamSynchronize(st));   // also: the host word copied to rowptr[na] is a local
    if (bad || 2 * (int64_t)P.n_upper != P.npairs + na)
        throw Error(CP_ESTATE, "star pair rows are not symmetric");
    P.urec.alloc(ctx, (size_t)P.n_upper * 16);
    launch(k_upper_rec, dim3(cdiv(P.n_upper, BLOCK)), dim3(BLOCK), 0, st, P.n_upper, (const int*)ulist.as<int>(),
           (const int*)rowof.as<int>(), (const int*)P.pb.as<int>(), (const int*)rev.as<int>(),
           (const int*)P.rowptr.as<int>(), P.urec.as<int4>());
    CP_HIP(hipStreamSynchronize(st));   // the local buffers above are released on return
}

static int bit_length(int64_t v) {
    int b = 1;
    while ((1LL << b) <= v) ++b;
    return b;
}

// Julia type-7 quantile.
static double quantile7(std::vector<double> v, double p) {
    std::sort(v.begin(), v.end());
    size_t n = v.size();
    double h = (n - 1) * p;
    size_t lo = (size_t)std::floor(h);
    if (lo + 1 >= n) return v[n - 1];
    return v[lo] + (h - lo) * (v[lo + 1] - v[lo]);
}

// _model_radii :auto (:209-243): the curve of growth is closed-form for the
// analytic models (curve_of_growth.jl:180-211) and pixel-integrated over exact
// circular apertures for ImagePSF / GriddedPSFModel (curve_of_growth.jl:143-176).
static std::vector<int> auto_radii(const HostPsf& P, const cp_psf_desc* psfd, double nsigma, int R_fit, int R_cap,
                                   const std::vector<double>& wsample, const std::vector<double>& flux) {
    std::vector<double> ws;
    for (double v : wsample)
        if (v > 0 && std::isfinite(v)) ws.push_back(v);
    CP_ARG(!ws.empty(), "no positive finite weights for the :auto noise estimate");
    double wv = quantile7(ws, 0.84);
    double sigma_bg = wv > 0 ? std::sqrt(1 / wv) : 1.0;
    std::vector<double> cg(R_cap);
    if (P.kind == CP_PSF_GRIDDED_IMAGE) {
        std::vector<double> radii(R_cap);
        for (int r = 1; r <= R_cap; ++r) radii[r - 1] = r;
        CP_ARG(model_cog_host(psfd, R_cap, radii.data(), cg.data()),
               "model_rad=:auto: the curve of growth of this PSF model could not be computed");
    } else {
        for (int r = 1; r <= R_cap; ++r)
            cg[r - 1] = P.kind == CP_PSF_CIRCULAR_GAUSSIAN
                            ? 1 - std::exp(CP_GAUSS_PRE * (double)r * r / (P.fwhm * P.fwhm))
                            : 1 - std::pow(1 + (double)r * r / (P.alpha * P.alpha), 1 - P.beta);
    }
    std::vector<double> sb(R_cap);
    double last = cg[R_cap - 1], prev = 0;
    for (int r = 1; r <= R_cap; ++r) {
        double ee = cg[r - 1] / last;
        sb[r - 1] = std::max((ee - prev) / (M_PI * (double)(r * r - (r - 1) * (r - 1))), 0.0);
        prev = ee;
    }
    double thr = nsigma * sigma_bg;
    std::vector<int> out(flux.size());
    for (size_t j = 0; j < flux.size(); ++j) {
        double F = std::max(flux[j], 0.0);
        int r = R_fit;
        while (r < R_cap && F * sb[r - 1] > thr) ++r;
        out[j] = r;
    }
    return out;
}


// ---------------------------------------------------------------------------
// the fit
// ---------------------------------------------------------------------------

// Instantiated HIP graphs of one fit's trial body (inner solve, step,
// candidate render and cost), keyed by the one argument that changes between
// trials (which of the two residual buffers is current).
struct TrialGraphs {
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    const void* key[2] = {nullptr, nullptr};
    ~TrialGraphs() {
        for (auto& e : exec)
            if (e) (void)hipGraphExecDestroy(e);
    }
    // Launch the graph for `key`, capturing `body` on `st` the first time.
    template <typename F>
    void run(hipStream_t st, const void* k, F&& body) {
        int slot = key[0] == k ? 0 : (key[1] == k ? 1 : (key[0] == nullptr ? 0 : 1));
        if (key[slot] != k) {
            if (exec[slot]) CP_HIP(hipGraphExecDestroy(exec[slot]));
            exec[slot] = nullptr;
            hipGraph_t g = nullptr;
            CP_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            try {
                body();
            } catch (...) {
                hipGraph_t junk = nullptr;
                (void)hipStreamEndCapture(st, &junk);
                if (junk) (void)hipGraphDestroy(junk);
                throw;
            }
            CP_HIP(hipStreamEndCapture(st, &g));
            hipError_t e = hipGraphInstantiate(&exec[slot], g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            CP_HIP(e);
            key[slot] = k;
        }
        CP_HIP(hipGraphLaunch(exec[slot], st));
    }
};
template <typename T, int KIND>
struct Sim {
    // storage precision of the equilibrated stamp Jacobian (J^T r and the
    // normal matrix of the inner solve are summed from the stored values, in
    // float64)
    using VT = T;
    cp_ctx* ctx;
    const HostPsf& P;
    DevPsf& DP;
    hipStream_t st;

    const cp_psf_desc* psfd_;
    bool fixed_q = false;   // the ePSF grid is staged with RQF x RQF planes
    int rt_ = RT;           // render tile side (cp_tuning.render_tile)
    int fill_gl_ = 16;      // lanes per star of the fill for stamps of <= 64 pixels (cp_tuning.fill_lanes)

    Sim(cp_ctx* c, const HostPsf& p, DevPsf& d, const cp_psf_desc* pd)
        : ctx(c), P(p), DP(d), st(c->stream), psfd_(pd) {}

    template <bool FINAL, bool LDSG>
    void fill_launch(int S2, dim3 grid, size_t sh, const ModelConst<T>& M, const StarRec<T>* recs,
                     const EpsfRec<T>* ers, const CornerRec<T>* crs, const int* lst, int na, int R, int S,
                     const FreeMap& fm, const T* w, int ny, int nx, const T* wres, VT* V, double* cn, double* b,
                     uint8_t* gloc, double gtol, double* part, Scal* Sc) {
        const char* nm = FINAL ? "k_fill_final" : "k_fill";
#define CP_FILL(E, G)                                                                                        \
    launch_n(nm, k_fill<T, KIND, FINAL, LDSG, E, G, VT>, grid, dim3(FBLOCK), sh, st, M, recs, ers, crs, lst, na, R, S, fm, \
             w, ny, nx, wres, V, cn, b, gloc, gtol, part, Sc)
        if (S2 <= 64 && fill_gl_ == 16) CP_FILL(4, 16);
        else if (S2 <= 64) CP_FILL(1, 64);
        else if (S2 <= 128) CP_FILL(2, 64);
        else if (S2 <= 256) CP_FILL(4, 64);
        else if (S2 <= 512) CP_FILL(8, 64);
        else CP_FILL(16, 64);
#undef CP_FILL
    }

    template <int MODE, int RTV>
    void render_launch_rt(bool ldsg, dim3 grid, size_t sh, const ModelConst<T>& M, const StarRec<T>* recs,
                          const EpsfRec<T>* ers, const CornerRec<T>* crs, const TileLists& L, const TileGeo& g,
                          int ntiles, const int2* shares, const uint8_t* umask, T* dw, const T* w, const T* data,
                          T* wres, T* resid, double* part, Scal* Sc, const int* tmark, int mark_iter) {
        static const char* nm[4] = {"k_render_resid", "k_render_cost", "k_render_freeze", "k_render_final"};
        if (ldsg && fixed_q)
            launch_n(nm[MODE], k_render<T, KIND, MODE, true, (KIND == KIND_EPSF ? RQF : 0), RTV>, grid, dim3(RBLOCK),
                     sh, st, M, recs, ers, crs, (const int*)L.start.as<int>(), (const int*)L.end.as<int>(),
                     (const int*)L.list.as<int>(), g, ntiles, shares, umask, dw, w, data, wres, resid, part, Sc, tmark,
                     mark_iter);
        else if (ldsg)
            launch_n(nm[MODE], k_render<T, KIND, MODE, true, 0, RTV>, grid, dim3(RBLOCK), sh, st, M, recs, ers, crs,
                     (const int*)L.start.as<int>(), (const int*)L.end.as<int>(), (const int*)L.list.as<int>(), g,
                     ntiles, shares, umask, dw, w, data, wres, resid, part, Sc, tmark, mark_iter);
        else
            launch_n(nm[MODE], k_render<T, KIND, MODE, false, 0, RTV>, grid, dim3(RBLOCK), sh, st, M, recs, ers, crs,
                     (const int*)L.start.as<int>(), (const int*)L.end.as<int>(), (const int*)L.list.as<int>(), g,
                     ntiles, shares, umask, dw, w, data, wres, resid, part, Sc, tmark, mark_iter);
    }

    template <int MODE>
    void render_launch(bool ldsg, dim3 grid, size_t sh, const ModelConst<T>& M, const StarRec<T>* recs,
                       const EpsfRec<T>* ers, const CornerRec<T>* crs, const TileLists& L, const TileGeo& g,
                       int ntiles, const int2* shares, const uint8_t* umask, T* dw, const T* w, const T* data,
                       T* wres, T* resid, double* part, Scal* Sc, const int* tmark = nullptr, int mark_iter = 0) {
        if (rt_ == 32)
            render_launch_rt<MODE, 32>(ldsg, grid, sh, M, recs, ers, crs, L, g, ntiles, shares, umask, dw, w, data,
                                       wres, resid, part, Sc, tmark, mark_iter);
        else
            render_launch_rt<MODE, 64>(ldsg, grid, sh, M, recs, ers, crs, L, g, ntiles, shares, umask, dw, w, data,
                                       wres, resid, part, Sc, tmark, mark_iter);
    }

    void run(const float* d_img, const float* d_iv, int ny, int nx, const double* d_y0, const double* d_x0,
             const double* d_f0, int n, double fit_rad, const cp_sim_opts& o, cp_sim_result* out,
             const std::vector<double>& hy0, const std::vector<double>& hx0, const std::vector<double>& hf0) {
        const int64_t Pn = (int64_t)ny * nx;
        const int R = (int)std::ceil(fit_rad), S = 2 * R + 1, S2 = S * S;
        CP_ARG(S2 <= 64 * MAX_EPL, "fit_rad too large for this build (ceil(fit_rad) <= 15)");
        // work-tile sizes of the context (cp_tuning; 0 = these defaults).  An
        // image of fewer 64-pixel tiles than CUs renders in 32-pixel tiles,
        // four times the blocks (a 512 x 512 tile: 3-7 % faster per fit; from
        // 1024 x 1024 on the 64-pixel tiles are as fast or faster)
        const int64_t nt64 = (int64_t)((ny + RT - 1) / RT) * ((nx + RT - 1) / RT);
        rt_ = ctx->tuning.render_tile > 0 ? ctx->tuning.render_tile : (nt64 < ctx->n_cu ? 32 : RT);
        fill_gl_ = ctx->tuning.fill_lanes > 0 ? ctx->tuning.fill_lanes : 16;
        // free set (y, x, flux in getproperties order)
        FreeMap fm{};
        int gcs[3] = {0, 0, 0}, np_ = 0;
        if (o.free_mask & CP_FREE_Y) gcs[np_++] = 1;
        if (o.free_mask & CP_FREE_X) gcs[np_++] = 2;
        if (o.free_mask & CP_FREE_FLUX) gcs[np_++] = 3;
        fm.p = np_;
        fm.gc0 = gcs[0];
        fm.gc1 = gcs[1];
        fm.gc2 = gcs[2];
        const int p = fm.p;
        fm.y_fixed = !(o.free_mask & CP_FREE_Y);
        fm.x_fixed = !(o.free_mask & CP_FREE_X);
        fm.f_fixed = !(o.free_mask & CP_FREE_FLUX);
        fm.fixed_y = o.fixed_y;
        fm.fixed_x = o.fixed_x;
        fm.fixed_f = o.fixed_flux;
        int k_y = -1, k_x = -1;
        for (int k = 0; k < p; ++k) {
            if (gcs[k] == 1) k_y = k;
            if (gcs[k] == 2) k_x = k;
        }
        const double model_rad_max = o.model_rad_max > 0 ? o.model_rad_max : 10 * fit_rad;
        const int n_cu = std::max(ctx->n_cu, 1);

        // --- outputs for the empty / all-failed cases
        auto write_initial = [&](const std::vector<uint8_t>& valid) {
            for (int i = 0; i < n; ++i) {
                out->y[i] = hy0[i];
                out->x[i] = hx0[i];
                out->flux[i] = hf0[i];
                out->bkg[i] = o.fixed_bkg;
                out->y_err[i] = fm.y_fixed ? 0.0 : NAN;
                out->x_err[i] = fm.x_fixed ? 0.0 : NAN;
                out->flux_err[i] = fm.f_fixed ? 0.0 : NAN;
                out->bkg_err[i] = 0;
                out->converged[i] = 0;
                out->valid[i] = valid[i];
                out->n_iter[i] = 0;
                if (out->chisq) out->chisq[i] = 0;
                for (double* d : {out->qfit, out->qfit_expected, out->qfit_z, out->crowding, out->spread_model,
                                  out->spread_model_err})
                    if (d) d[i] = NAN;
            }
        };
        out->trace_len = 0;
        out->n_trials_total = 0;
        out->n_inner_total = 0;
        out->n_active = 0;
        out->n_union = 0;
        out->cost_final = 0;
        if (n == 0) {
            out->n_passes = 0;
            out->n_failed = 0;
            return;
        }

        // --- device scalars
        DevBuf dS(ctx, sizeof(Scal));
        Scal* Sc = dS.as<Scal>();
        Scal* hS = static_cast<Scal*>(pinned(ctx, sizeof(Scal)));
        CP_HIP(hipMemsetAsync(Sc, 0, sizeof(Scal), st));
        // the scalars reach the host through a one-block copy kernel into the
        // pinned mirror (cheaper than a copy-engine transfer of a few hundred bytes)
        auto read_S = [&]() {
            launch(k_mirror, dim3(1), dim3(64), 0, st, (const Scal*)Sc, hS);
            CP_HIP(hipStreamSynchronize(st));
        };

        // --- data and weights
        DevBuf data(ctx, Pn * sizeof(T)), dw(ctx, Pn * sizeof(T)), w(ctx, Pn * sizeof(T));
        launch_n("k_prep_pixels", k_prep_pixels<T>, dim3(std::min<int64_t>(cdiv(Pn, BLOCK), n_cu * 16)), dim3(BLOCK), 0, st, d_img,
               d_iv, Pn, data.as<T>(), dw.as<T>(), w.as<T>(), Sc);
        read_S();
        CP_ARG(hS->n_nonfinite == 0 || d_iv != nullptr,
               "the image contains non-finite values but `inv_var` was not provided; pass `inv_var` to mask "
               "them or clean the image first");

        // --- stamps, active list (Morton order), union
        Cub cub{ctx, DevBuf()};
        DevBuf ay_cat(ctx, (size_t)n * 4), ax_cat(ctx, (size_t)n * 4), ok(ctx, (size_t)n);
        launch(k_anchor<T>, dim3(cdiv(n, BLOCK)), dim3(BLOCK), 0, st, d_y0, d_x0, d_f0, n, (const T*)w.as<T>(), ny,
               nx, R, ay_cat.as<int>(), ax_cat.as<int>(), ok.as<uint8_t>());
        DevBuf act;
        int coord_bits = bit_length((int64_t)std::max(ny, nx) + R + 1);
        const int na = build_active(ctx, cub, ay_cat.as<int>(), ax_cat.as<int>(), ok.as<uint8_t>(), n, coord_bits, act);
        std::vector<int> hact(std::max(na, 1));
        if (na > 0) {
            CP_HIP(hipMemcpyAsync(hact.data(), act.get(), (size_t)na * 4, hipMemcpyDeviceToHost, st));
            CP_HIP(hipStreamSynchronize(st));
        }
        std::vector<uint8_t> valid(n, 0);
        for (int j = 0; j < na; ++j) valid[hact[j]] = 1;
        if (out->active)
            for (int i = 0; i < n; ++i) out->active[i] = valid[i];
        out->n_failed = n - na;
        if (na == 0) {
            write_initial(valid);
            if (out->residual) std::memset(out->residual, 0, sizeof(float) * Pn);
            out->n_passes = 0;
            return;
        }
        DevBuf umask(ctx, Pn);
        CP_HIP(hipMemsetAsync(umask.get(), 0, Pn, st));
        DevBuf ay(ctx, (size_t)na * 4), ax(ctx, (size_t)na * 4), mR(ctx, (size_t)na * 4);
        DevBuf theta(ctx, (size_t)na * p * 8), live(ctx, na), frozen_at(ctx, (size_t)na * 4), conv(ctx, na);

        // --- model radii (:1113-1124)
        CP_ARG(!(o.model_rad <= 0 && d_iv == nullptr),
               "model_rad=:auto requires inv_var (it needs a noise scale); pass inv_var or an explicit scalar model_rad");
        const int R_cap = (int)std::ceil(o.model_rad <= 0 ? model_rad_max : std::max(o.model_rad, fit_rad));
        int mR_scalar = 0;
        DevBuf mR_src;
        if (o.model_rad > 0) {
            mR_scalar = std::min(std::max((int)std::ceil(o.model_rad), R), R_cap);
        } else {
            int64_t step = std::max<int64_t>(1, Pn / 65536);
            int64_t ns = (Pn + step - 1) / step;
            std::vector<T> hw(ns);
            // strided sample of w (w[1:step:end])
            DevBuf samp(ctx, ns * sizeof(T));
            CP_HIP(hipMemcpy2DAsync(samp.get(), sizeof(T), w.get(), step * sizeof(T), sizeof(T), ns,
                                    hipMemcpyDeviceToDevice, st));
            CP_HIP(hipMemcpyAsync(hw.data(), samp.get(), ns * sizeof(T), hipMemcpyDeviceToHost, st));
            CP_HIP(hipStreamSynchronize(st));
            std::vector<double> wsd(hw.begin(), hw.end()), fl(na);
            for (int j = 0; j < na; ++j) fl[j] = hf0[hact[j]];
            std::vector<int> rr = auto_radii(P, psfd_, o.model_rad_nsigma, R, R_cap, wsd, fl);
            mR_src.alloc(ctx, (size_t)na * 4);
            CP_HIP(hipMemcpyAsync(mR_src.get(), rr.data(), (size_t)na * 4, hipMemcpyHostToDevice, st));
            CP_HIP(hipStreamSynchronize(st));
        }
        launch(k_star_setup, dim3(cdiv(na, BLOCK)), dim3(BLOCK), 0, st, (const int*)act.as<int>(), na,
               (const int*)ay_cat.as<int>(), (const int*)ax_cat.as<int>(), d_y0, d_x0, d_f0, p, fm.gc0, fm.gc1,
               fm.gc2, (const int*)(o.model_rad > 0 ? nullptr : mR_src.as<int>()), mR_scalar, ay.as<int>(),
               ax.as<int>(), mR.as<int>(), theta.as<double>(), live.as<uint8_t>(), frozen_at.as<int>(),
               conv.as<uint8_t>());
        launch_n("k_union", k_union<T>, dim3(cdiv((int64_t)na * S2, BLOCK)), dim3(BLOCK), 0, st, na, (const int*)ay.as<int>(),
               (const int*)ax.as<int>(), R, S, (const T*)w.as<T>(), ny, nx, umask.as<uint8_t>());
        {
            const int gcm = (int)std::max<int64_t>(1, std::min<int64_t>(cdiv(Pn / 4, BLOCK), n_cu * 4));
            DevBuf cm_part(ctx, (size_t)gcm * 8);
            launch(k_count_mask, dim3(gcm), dim3(BLOCK), 0, st, (const uint8_t*)umask.as<uint8_t>(), Pn,
                   cm_part.as<double>());
            launch(k_finish<FIN_COUNT>, dim3(1), dim3(FIN_BLOCK), 0, st, Sc, (const double*)cm_part.as<double>(),
                   (const double*)nullptr, gcm, 0, (Scal*)nullptr);
            read_S();
        }
        const int64_t U = (int64_t)hS->n_union;
        const int64_t nvar = (int64_t)p * na;
        const double dof = (double)std::max<int64_t>(U - nvar, 1);
        CP_HIP(hipMemcpyAsync(&Sc->dof, &dof, 8, hipMemcpyHostToDevice, st));

        // --- model-box tile lists (render passes, rt x rt tiles) and the rows
        // of star pairs with overlapping stamps (the normal matrix H = J^T J)
        const int rt = rt_;
        TileGeo g{rt, rt, (ny + rt - 1) / rt, (nx + rt - 1) / rt, ny, nx};
        const int ntiles = g.nty * g.ntx;
        TileLists Lmod;
        PairRows prs;
        build_pairs(ctx, cub, na, ay.as<int>(), ax.as<int>(), R, ny, nx, prs);
        build_tiles(ctx, cub, na, ay.as<int>(), ax.as<int>(), o.model_rad > 0 ? nullptr : mR.as<int>(),
                    mR_scalar, g, Lmod);
        // cost-balanced tile shares of the render blocks
        DevBuf tcost(ctx, (size_t)ntiles * 8), tscan(ctx, (size_t)ntiles * 8);
        launch(k_tile_cost, dim3(cdiv(ntiles, BLOCK)), dim3(BLOCK), 0, st, (const int*)Lmod.start.as<int>(),
               (const int*)Lmod.end.as<int>(), ntiles, tcost.as<long long>());
        {
            size_t b = 0;
            CP_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, b, tcost.as<long long>(), tscan.as<long long>(), ntiles, st));
            cub.ensure(b);
            b = cub.tmp.bytes();
            CP_HIP(hipcub::DeviceScan::InclusiveSum(cub.tmp.get(), b, tcost.as<long long>(), tscan.as<long long>(),
                                                    ntiles, st));
        }
        cub.tmp.release();

        // --- working buffers
        // weighted residual of the current point and of the trial candidate
        DevBuf wres_a(ctx, Pn * sizeof(T)), wres_b(ctx, Pn * sizeof(T)), ubuf(ctx, Pn * sizeof(T));
        CP_HIP(hipMemsetAsync(wres_a.get(), 0, Pn * sizeof(T), st));
        CP_HIP(hipMemsetAsync(wres_b.get(), 0, Pn * sizeof(T), st));
        T* wres_cur = wres_a.as<T>();
        T* wres_cand = wres_b.as<T>();
        CP_HIP(hipMemsetAsync(ubuf.get(), 0, Pn * sizeof(T), st));
        DevBuf V(ctx, (size_t)na * S2 * VP * sizeof(VT));
        CP_HIP(hipMemsetAsync(V.get(), 0, V.bytes(), st));
        const size_t nb8 = (size_t)na * p * 8;
        const size_t nb32 = (size_t)na * VP * 8;   // b_scaled: padded like the V rows it is summed from
        const int solver = o.solver;
        DevBuf colnorm(ctx, nb8), D(ctx, nb8), bsc(ctx, nb32), xs(ctx, nb8), wl(ctx, nb8), dsc(ctx, nb8),
            theta_c(ctx, nb8), hbar(ctx, solver == CP_SOLVER_LSMR ? nb8 : 0);
        // the normal matrix's p x p blocks and two pairs of star vectors:
        // LSQR / LSMR (g, v) of the previous and the current Golub-Kahan step,
        // CG s, q = H p and the directions of two successive steps
        DevBuf Hblk(ctx, (size_t)std::max<int64_t>(prs.npairs, 1) * p * p * 8), Hdiag(ctx, (size_t)na * p * p * 8),
            gv0(ctx, nb8), gv1(ctx, nb8), vv0(ctx, nb8), vv1(ctx, nb8);
        // the entry words of the rows live at the last compaction, packed (cent),
        // each entry's packed position (cpos, -1 for a frozen row) and each
        // live row's start (cst), from the live rows' lengths (rlen)
        const size_t ne_all = (size_t)std::max<int64_t>(prs.npairs, 1);
        DevBuf cent(ctx, ne_all * 8), cpos(ctx, ne_all * 4), cst(ctx, (size_t)(na + 1) * 4),
            rlen(ctx, (size_t)(na + 1) * 4);
        for (DevBuf* b : {&gv0, &gv1, &vv0, &vv1, &colnorm, &xs, &wl}) CP_HIP(hipMemsetAsync(b->get(), 0, nb8, st));
        CP_HIP(hipMemsetAsync(bsc.get(), 0, nb32, st));
        DevBuf gloc(ctx, na), recs(ctx, (size_t)na * sizeof(StarRec<T>)), recs_c(ctx, (size_t)na * sizeof(StarRec<T>));
        DevBuf crs, crs_c, ers, ers_c;
        const bool multi = (KIND == KIND_EPSF && P.n_nodes > 1);
        const bool single_epsf = (KIND == KIND_EPSF && P.n_nodes == 1);
        ers.alloc(ctx, (size_t)(single_epsf ? na : 1) * sizeof(EpsfRec<T>));
        ers_c.alloc(ctx, (size_t)(single_epsf ? na : 1) * sizeof(EpsfRec<T>));
        if (multi) {
            crs.alloc(ctx, (size_t)na * sizeof(CornerRec<T>));
            crs_c.alloc(ctx, (size_t)na * sizeof(CornerRec<T>));
        } else {
            crs.alloc(ctx, sizeof(CornerRec<T>));
            crs_c.alloc(ctx, sizeof(CornerRec<T>));
        }

        ModelConst<T> M = model_const<T>(P, DP, o.fixed_bkg);
        M.poly = nullptr;
        const double inv_norm_g64 = (KIND == KIND_GAUSS) ? 1.0 / (-(M_PI * P.fwhm * P.fwhm / CP_GAUSS_PRE)) : 0.0;
        const double moffat_amp64 = (KIND == KIND_MOFFAT) ? (P.beta - 1) / (M_PI * P.alpha * P.alpha) : 0.0;
        // os 4 grids whose planes fit RQF x RQF cells are staged with planes of
        // that side (the renders' fixed-layout path)
        fixed_q = single_epsf && M.ios_y == 4 && M.ios_x == 4 && M.Qy <= RQF && M.Qx <= RQF;
        if (fixed_q) M.Qy = M.Qx = RQF;
        const size_t poly_words = single_epsf ? (size_t)M.ios_y * M.ios_x * M.Qy * M.Qx : 0;
        // the LDS grid copy (edge-padded, polyphase) leaves room for two float
        // render / fill blocks per CU (one for float64)
        const bool ldsg = single_epsf && poly_words * 4 <= (sizeof(T) == 4 ? 56 : 96) * 1024;
        const size_t grid_sh = ldsg ? ((poly_words * 4 + 15) & ~(size_t)15) : 0;
        const size_t render_sh = render_lds_bytes<T>(ldsg ? (int)poly_words : 0, rt_);
        DevBuf polyg;
        if (ldsg) {
            polyg.alloc(ctx, poly_words * 4);
            CP_HIP(hipMemsetAsync(polyg.get(), 0, poly_words * 4, st));
            launch(k_poly_grid, dim3(cdiv((int64_t)poly_words, BLOCK)), dim3(BLOCK), 0, st, (const float*)M.grids, M.gny,
                   M.gnx, M.ios_y, M.ios_x, M.Qy, M.Qx, polyg.as<float>());
            M.poly = polyg.as<float>();
        }
        CP_ARG(render_sh <= 160 * 1024, "render LDS footprint above 160 KiB");
        const size_t fill_sh = grid_sh;

        // two 1024-thread render / fill blocks fit one CU (LDS and wave slots);
        // render blocks take equal shares of the tile-cost scan
        // (blocks per CU: the context's tuned launch shapes, else these defaults)
        const cp_tuning& tu = ctx->tuning;
        // (two 1024-thread blocks per CU: 1-2.5 % faster than four on 2048^2 and
        // 4096^2 tiles, the same on smaller ones)
        const int gb_render = (int)std::min<int64_t>(ntiles, (int64_t)n_cu * tuned_bpc(tu.render_blocks_per_cu, 2));
        // each render block's tiles over all stars' lists (Lmod) and the live
        // lists (rebuilt with them)
        DevBuf shares_m(ctx, (size_t)gb_render * sizeof(int2)), shares_l(ctx, (size_t)gb_render * sizeof(int2));
        launch(k_tile_shares, dim3(cdiv(gb_render, BLOCK)), dim3(BLOCK), 0, st, (const long long*)tscan.as<long long>(),
               ntiles, gb_render, shares_m.as<int2>());
        // CG's H p: GKL lanes per row, a group for
        // every star; CG's update: one thread per star
        const int gb_jv = (int)std::max<int64_t>(1, cdiv(na, BLOCK / GKL));
        const int gb_jtu = (int)std::max<int64_t>(
            1, std::min<int64_t>(cdiv(na, BLOCK), (int64_t)n_cu * tuned_bpc(tu.cg_blocks_per_cu, 8)));
        // the H assembly: GHh lanes per stored block (by stamp size unless set)
        const int GHh = tu.hbuild_lanes > 0 ? tu.hbuild_lanes : (S2 <= 64 ? 4 : (S2 <= 256 ? 16 : 64));
        const int gb_hb = (int)std::max<int64_t>(
            1, std::min<int64_t>(cdiv(prs.n_upper, BLOCK / GHh), (int64_t)n_cu * tuned_bpc(tu.hbuild_blocks_per_cu, 32)));
        const int gb_fill = (int)std::min<int64_t>(cdiv(na, FBLOCK / 64), (int64_t)n_cu * 2);
        const unsigned int gstar = cdiv(na, BLOCK);
        // per-block partials of the widest launch (two sets for fill / accept)
        const int maxpb = (int)std::max<int64_t>(
            {(int64_t)gb_render, (int64_t)gb_fill, (int64_t)gb_jv, (int64_t)gb_jtu, (int64_t)gstar});
        DevBuf part(ctx, (size_t)2 * maxpb * 8), part2(ctx, (size_t)maxpb * 8);
        // CG's partials (two per block)
        DevBuf partk(ctx, (size_t)2 * std::max(gb_jv, gb_jtu) * 8);
        // the Golub-Kahan step: one lane per row entry; a block sums the rows
        // headed in its first gk_span entries, the rest of its lanes cover rows
        // that run past them (up to the longest row, at most half the block)
        const int gk_margin = std::min(GKB / 2, std::max(64, 64 * (int)cdiv((int64_t)prs.max_row - 1, 64)));
        const int gk_span = GKB - gk_margin;
        const int gb_gk = (int)std::max<int64_t>(1, cdiv(prs.npairs, gk_span));
        // grids of at most one block per CU fold each step's reduction into
        // the next step (k_gk<..., true>): its partials and scalar state
        // alternate between two buffers
        const bool gk_fold = gb_gk <= (tu.gk_fold_blocks > 0 ? tu.gk_fold_blocks : n_cu);
        // its partials (ten per block for LSMR, five for LSQR)
        DevBuf partg(ctx, (size_t)(gk_fold ? 20 : 10) * gb_gk * 8), scal_gk(ctx, 2 * sizeof(Scal));

        // star records of the stars in lst (null: all)
        auto star_rec = [&](const int* lst, const double* th, DevBuf& r, DevBuf& c, DevBuf& e, int frozen_iter,
                            int all_live) {
            launch_n("k_star_rec", k_star_rec<T, KIND>, dim3(gstar), dim3(BLOCK), 0, st, na, lst, (const Scal*)Sc, th, fm, M,
                   inv_norm_g64, moffat_amp64,
                   (const int*)ay.as<int>(), (const int*)ax.as<int>(), (const int*)mR.as<int>(),
                   (const uint8_t*)live.as<uint8_t>(), (const int*)frozen_at.as<int>(), frozen_iter, all_live,
                   r.as<StarRec<T>>(), DP.G, c.as<CornerRec<T>>(), single_epsf ? e.as<EpsfRec<T>>() : nullptr);
        };
        // live stars of the last compaction (fill and J^T u walk them) and the
        // model-box tile lists restricted to them (cost and freeze renders)
        DevBuf jlist(ctx, (size_t)na * 4);
        TileLists Llive;
        Llive.npairs = Lmod.npairs;
        Llive.start.alloc(ctx, (size_t)ntiles * 4);
        Llive.end.alloc(ctx, (size_t)ntiles * 4);
        Llive.list.alloc(ctx, (size_t)std::max(Lmod.npairs, 1) * 4);
        DevBuf tflag(ctx, (size_t)(Lmod.npairs + 1) * 4), tpos(ctx, (size_t)(Lmod.npairs + 1) * 4);
        DevBuf tcost_l(ctx, (size_t)ntiles * 8), tscan_l(ctx, (size_t)ntiles * 8);
        // per tile: the last iteration at which a star touching it froze
        DevBuf tmark(ctx, (size_t)ntiles * 4);
        CP_HIP(hipMemsetAsync(tmark.get(), 0, (size_t)ntiles * 4, st));
        auto fill = [&](bool final_mode) {
            const StarRec<T>* rp = recs.as<StarRec<T>>();
            const EpsfRec<T>* ep = ers.as<EpsfRec<T>>();
            const CornerRec<T>* cp = crs.as<CornerRec<T>>();
            if (ldsg) {
                if (final_mode)
                    fill_launch<true, true>(S2, dim3(gb_fill), fill_sh, M, rp, ep, cp, nullptr, na, R, S, fm, w.as<T>(), ny, nx, (const T*)wres_cur, V.as<VT>(), colnorm.as<double>(), bsc.as<double>(), gloc.as<uint8_t>(), o.g_tol, part.as<double>(), Sc);
                else
                    fill_launch<false, true>(S2, dim3(gb_fill), fill_sh, M, rp, ep, cp, (const int*)jlist.as<int>(), na, R, S, fm, w.as<T>(), ny, nx, (const T*)wres_cur, V.as<VT>(), colnorm.as<double>(), bsc.as<double>(), gloc.as<uint8_t>(), o.g_tol, part.as<double>(), Sc);
            } else {
                if (final_mode)
                    fill_launch<true, false>(S2, dim3(gb_fill), 0, M, rp, ep, cp, nullptr, na, R, S, fm, w.as<T>(), ny, nx, (const T*)wres_cur, V.as<VT>(), colnorm.as<double>(), bsc.as<double>(), gloc.as<uint8_t>(), o.g_tol, part.as<double>(), Sc);
                else
                    fill_launch<false, false>(S2, dim3(gb_fill), 0, M, rp, ep, cp, (const int*)jlist.as<int>(), na, R, S, fm, w.as<T>(), ny, nx, (const T*)wres_cur, V.as<VT>(), colnorm.as<double>(), bsc.as<double>(), gloc.as<uint8_t>(), o.g_tol, part.as<double>(), Sc);
            }
        };
        auto finish = [&](auto kern, int nb, int it, Scal* mirror = nullptr) {
            launch(kern, dim3(1), dim3(FIN_BLOCK), 0, st, Sc, (const double*)part.as<double>(),
                   (const double*)part2.as<double>(), nb, it, mirror);
        };

        // --- the stars live at the last compaction (the fill, the star
        // records and the freeze marks walk them; count into Sc) and the model-box tile lists
        // restricted to them (cost and freeze renders), rebuilt whenever the
        // live set has shrunk by a fifth.  The buffers never move, so
        // captured trial graphs keep their arguments.
        int live_compacted = -1;
        auto compact_live = [&]() {
            {
                launch(k_live_rowlen, dim3(cdiv(na + 1, BLOCK)), dim3(BLOCK), 0, st, na,
                       (const uint8_t*)live.as<uint8_t>(), (const int*)prs.rowptr.as<int>(), rlen.as<int>());
                size_t b = 0;
                CP_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, b, rlen.as<int>(), cst.as<int>(), na + 1, st));
                cub.ensure(b);
                b = cub.tmp.bytes();
                CP_HIP(hipcub::DeviceScan::ExclusiveSum(cub.tmp.get(), b, rlen.as<int>(), cst.as<int>(), na + 1, st));
                launch(k_live_cpos, dim3(cdiv(na, BLOCK)), dim3(BLOCK), 0, st, na, (const uint8_t*)live.as<uint8_t>(),
                       (const int*)prs.rowptr.as<int>(), (const int*)cst.as<int>(), cpos.as<int>(), Sc);
            }
            {
                hipcub::CountingInputIterator<int> cit(0);
                size_t b = 0;
                CP_HIP(hipcub::DeviceSelect::Flagged(nullptr, b, cit, (const uint8_t*)live.as<uint8_t>(), jlist.as<int>(),
                                                     &Sc->n_jtu_stars, na, st));
                cub.ensure(b);
                b = cub.tmp.bytes();
                CP_HIP(hipcub::DeviceSelect::Flagged(cub.tmp.get(), b, cit, (const uint8_t*)live.as<uint8_t>(),
                                                     jlist.as<int>(), &Sc->n_jtu_stars, na, st));
            }
            const int np = Lmod.npairs;
            const int gbl = (int)std::max<int64_t>(1, std::min<int64_t>(cdiv((int64_t)np + 1, BLOCK), (int64_t)n_cu * 16));
            launch(k_list_live, dim3(gbl), dim3(BLOCK), 0, st, np, (const int*)Lmod.list.as<int>(),
                   (const uint8_t*)live.as<uint8_t>(), tflag.as<int>());
            size_t b2 = 0;
            CP_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, b2, tflag.as<int>(), tpos.as<int>(), np + 1, st));
            cub.ensure(b2);
            b2 = cub.tmp.bytes();
            CP_HIP(hipcub::DeviceScan::ExclusiveSum(cub.tmp.get(), b2, tflag.as<int>(), tpos.as<int>(), np + 1, st));
            launch(k_list_compact, dim3(gbl), dim3(BLOCK), 0, st, np, (const int*)Lmod.list.as<int>(),
                   (const int*)tflag.as<int>(), (const int*)tpos.as<int>(), Llive.list.as<int>(), ntiles,
                   (const int*)Lmod.start.as<int>(), (const int*)Lmod.end.as<int>(), Llive.start.as<int>(),
                   Llive.end.as<int>());
            launch(k_tile_cost, dim3(cdiv(ntiles, BLOCK)), dim3(BLOCK), 0, st, (const int*)Llive.start.as<int>(),
                   (const int*)Llive.end.as<int>(), ntiles, tcost_l.as<long long>());
            b2 = 0;
            CP_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, b2, tcost_l.as<long long>(), tscan_l.as<long long>(),
                                                    ntiles, st));
            cub.ensure(b2);
            b2 = cub.tmp.bytes();
            CP_HIP(hipcub::DeviceScan::InclusiveSum(cub.tmp.get(), b2, tcost_l.as<long long>(),
                                                    tscan_l.as<long long>(), ntiles, st));
            launch(k_tile_shares, dim3(cdiv(gb_render, BLOCK)), dim3(BLOCK), 0, st,
                   (const long long*)tscan_l.as<long long>(), ntiles, gb_render, shares_l.as<int2>());
        };
        auto build_H = [&]() {
            with_p(p, [&](auto pc) {
                constexpr int PV = decltype(pc)::value;
#define CP_HB(GH)                                                                                                  \
    launch_n("k_hbuild", k_hbuild<VT, GH, PV>, dim3(gb_hb), dim3(BLOCK), 0, st, (const int4*)prs.urec.as<int4>(),    \
             prs.n_upper, (const uint8_t*)live.as<uint8_t>(), (const int*)ay.as<int>(), (const int*)ax.as<int>(), R,   \
             S, (const VT*)V.as<VT>(), Hblk.as<double>(), Hdiag.as<double>(), (const int*)cpos.as<int>(),         \
             cent.as<int2>())
                if (GHh == 4) CP_HB(4);
                else if (GHh == 16) CP_HB(16);
                else CP_HB(64);
#undef CP_HB
            });
        };

        // --- outer damped Gauss-Newton loop (:1186-1284)
        // An accepted candidate's weighted residual (written by its cost
        // render) is the next iteration's residual: the live-star model minus
        // the frozen-star-subtracted data is the same image, so only the first
        // iteration renders it separately.
        //
        // The host reads the device scalars once per damping trial (after its
        // candidate cost).  An accepted step's freezing is not read back: the
        // freeze render follows every accepted step and does nothing where no
        // star froze, and the live count after the step reaches the host with
        // the next iteration's first trial, so the live lists are recompacted
        // one iteration after the live set has shrunk by a fifth, and an
        // iteration that starts with no live star is discarded there.
        double lam = o.lambda_init, cost = 0;
        int n_run = 0, n_live_now = na;
        bool have_live_count = false;   // Sc->n_live is the count after an accepted step
        TrialGraphs graphs;
        const bool use_graphs = !ctx->prof;   // per-launch event timing needs plain launches
        bool D_init = false, have_resid = false, recs_valid = false;
        for (int iter = 1; iter <= o.max_iter; ++iter) {
            if (live_compacted < 0 || n_live_now < 0.8 * live_compacted) {
                compact_live();
                live_compacted = n_live_now;
            }
            // the star records of theta: built once; afterwards an accepted
            // step hands over its candidate's (theta_c is the new theta),
            // rebuilt where stars froze, and a rejected step leaves them
            if (!recs_valid) {
                star_rec(jlist.as<int>(), theta.as<double>(), recs, crs, ers, 0, 0);
                recs_valid = true;
            }
            if (!have_resid) {
                render_launch<MODE_RESID>(ldsg, dim3(gb_render), render_sh, M, recs.as<StarRec<T>>(),
                                          ers.as<EpsfRec<T>>(), crs.as<CornerRec<T>>(), Lmod, g, ntiles, (const int2*)shares_m.as<int2>(),
                                          umask.as<uint8_t>(), dw.as<T>(), w.as<T>(), data.as<T>(), wres_cur, nullptr,
                                          part.as<double>(), Sc);
                finish(k_finish<FIN_COST>, gb_render, 0);
                have_resid = true;
            }
            fill(false);   // its partials are summed by the first trial's k_lsqr_init
            build_H();
            if (!D_init) {
                CP_HIP(hipMemcpyAsync(D.get(), colnorm.get(), nb8, hipMemcpyDeviceToDevice, st));
                D_init = true;
            }
            // The cost and the gradient cosine of this iteration are read with
            // the first trial's result: the trial is launched without waiting
            // for them and discarded when the g test stops the fit (a trial
            // changes none of the state the stop keeps: theta, the residual,
            // the live set).
            double gnorm = 0;
            bool g_stop = false, none_live = false;
            bool accepted = false, f_conv = false;
            for (int trial = 1; trial <= o.max_trials; ++trial) {
                const double lam_solve = std::sqrt(lam);
                launch(k_lsqr_init, dim3(1), dim3(FIN_BLOCK), 0, st, Sc, lam_solve, o.inner_iterations, o.linear_tol,
                       o.linear_tol, solver, trial == 1 ? (const double*)part.as<double>() : (const double*)nullptr,
                       gb_fill);
                // inner solve, capped step and candidate cost: one graph launch per trial
                auto trial_body = [&]() {
                    with_p(p, [&](auto pc) {
                        constexpr int PV = decltype(pc)::value;
                        const int* rp = (const int*)prs.rowptr.as<int>();
                        const int2* en = (const int2*)cent.as<int2>();
                        const double* Hdp = (const double*)Hdiag.as<double>();
                        const int* jl = (const int*)jlist.as<int>();
                        const uint8_t* lv = (const uint8_t*)live.as<uint8_t>();
                        const double* Hp = (const double*)Hblk.as<double>();
                        const double* bs = (const double*)bsc.as<double>();
                        for (int it = 1; it <= o.inner_iterations; ++it) {
                            const int first = it == 1;
                            if (solver == CP_SOLVER_CGLS) {
                                // s in gv0, q in gv1; this step's direction goes to
                                // pcur, the previous one is read from pold
                                double* pcur = (it & 1) ? vv1.as<double>() : vv0.as<double>();
                                double* pold = (it & 1) ? vv0.as<double>() : vv1.as<double>();
                                launch_n("k_cg_hv", k_cg_hv<PV>, dim3(gb_jv), dim3(BLOCK), 0, st, na, lv, rp,
                                         (const int*)cst.as<int>(), en, Hp, Hdp, bs,
                                         (const double*)gv0.as<double>(), (const double*)pold, pcur, gv1.as<double>(),
                                         first, partk.as<double>(), Sc);
                                launch(k_finish<FIN_CGQ>, dim3(1), dim3(FIN_BLOCK), 0, st, Sc,
                                       (const double*)partk.as<double>(), (const double*)partk.as<double>() + gb_jv,
                                       gb_jv, it, (Scal*)nullptr);
                                launch_n("k_cg_update", k_cg_update<PV>, dim3(gb_jtu), dim3(BLOCK), 0, st, jl, lv, bs,
                                         (const double*)pcur, (const double*)gv1.as<double>(), gv0.as<double>(),
                                         xs.as<double>(), first, partk.as<double>(), Sc);
                                launch(k_finish<FIN_CGS>, dim3(1), dim3(FIN_BLOCK), 0, st, Sc,
                                       (const double*)partk.as<double>(), (const double*)partk.as<double>() + gb_jtu,
                                       gb_jtu, it, (Scal*)nullptr);
                                continue;
                            }
                            // (g, v) of step it - 1 in buffer set (it - 1) & 1, step it writes set it & 1
                            DevBuf& gprev = (it & 1) ? gv0 : gv1;
                            DevBuf& vprev = (it & 1) ? vv0 : vv1;
                            DevBuf& gnew = (it & 1) ? gv1 : gv0;
                            DevBuf& vnew = (it & 1) ? vv1 : vv0;
                            auto gk = [&](auto sv) {
                                constexpr int SV = decltype(sv)::value;
                                if (gk_fold) {
                                    // step it: partials to set it & 1, the state before step it - 1's
                                    // reduction from scal_gk[it & 1] (the fit's own for the first step),
                                    // after it to scal_gk[(it + 1) & 1]
                                    double* pc = partg.as<double>() + (size_t)(it & 1) * 10 * gb_gk;
                                    const double* pp = partg.as<double>() + (size_t)((it - 1) & 1) * 10 * gb_gk;
                                    Scal* sg = scal_gk.as<Scal>();
                                    const Scal* sin = first ? (const Scal*)Sc : sg + (it & 1);
                                    Scal* snext = sg + ((it + 1) & 1);
                                    launch_n("k_gk", k_gk<PV, SV, true>, dim3(gb_gk), dim3(GKB), 0, st, gk_span, en,
                                             Hp, Hdp, bs, (const double*)gprev.as<double>(),
                                             (const double*)vprev.as<double>(), gnew.as<double>(),
                                             vnew.as<double>(), xs.as<double>(), wl.as<double>(), hbar.as<double>(),
                                             first, pc, sin, pp, snext, it);
                                    // (the last step's reduction closes the solve in k_step_fold)
                                    return;
                                }
                                launch_n("k_gk", k_gk<PV, SV, false>, dim3(gb_gk), dim3(GKB), 0, st, gk_span, en, Hp,
                                         Hdp, bs, (const double*)gprev.as<double>(), (const double*)vprev.as<double>(),
                                         gnew.as<double>(), vnew.as<double>(), xs.as<double>(), wl.as<double>(),
                                         hbar.as<double>(), first, partg.as<double>(), (const Scal*)Sc,
                                         (const double*)nullptr, (Scal*)nullptr, it);
                                launch_n("k_finish_gk", k_finish_gk<SV>, dim3(1), dim3(FINGK_BLOCK), 0, st,
                                         (const Scal*)Sc, Sc, (const double*)partg.as<double>(), gb_gk, it);
                            };
                            if (solver == CP_SOLVER_LSMR) gk(std::integral_constant<int, CP_SOLVER_LSMR>{});
                            else gk(std::integral_constant<int, CP_SOLVER_LSQR>{});
                        }
                    });
                    const double* hbp = solver == CP_SOLVER_LSMR ? hbar.as<double>() : nullptr;
                    if (gk_fold && solver != CP_SOLVER_CGLS) {
                        // the last Golub-Kahan step's reduction (its partials in set
                        // inner & 1, the state before it in scal_gk[(inner + 1) & 1])
                        // with the step itself
                        const int itm = o.inner_iterations;
                        const double* pc = partg.as<double>() + (size_t)(itm & 1) * 10 * gb_gk;
                        const Scal* sin = scal_gk.as<Scal>() + ((itm + 1) & 1);
                        auto sf = [&](auto sv) {
                            constexpr int SV = decltype(sv)::value;
                            launch_n("k_step", k_step_fold<SV>, dim3(cdiv(na, FINGK_BLOCK)), dim3(FINGK_BLOCK), 0, st,
                                     na, p, k_y, k_x, o.max_step, (const double*)wl.as<double>(), hbp,
                                     xs.as<double>(), (const double*)colnorm.as<double>(),
                                     (const double*)theta.as<double>(), dsc.as<double>(), theta_c.as<double>(),
                                     (const uint8_t*)live.as<uint8_t>(), sin, Sc, pc, gb_gk, itm);
                        };
                        if (solver == CP_SOLVER_LSMR) sf(std::integral_constant<int, CP_SOLVER_LSMR>{});
                        else sf(std::integral_constant<int, CP_SOLVER_LSQR>{});
                    } else {
                        launch_n("k_step", k_step, dim3(gstar), dim3(BLOCK), 0, st, na, p, k_y, k_x, o.max_step,
                                 (const double*)wl.as<double>(), hbp, xs.as<double>(),
                                 (const double*)colnorm.as<double>(), (const double*)theta.as<double>(),
                                 dsc.as<double>(), theta_c.as<double>(), (const uint8_t*)live.as<uint8_t>(),
                                 (const Scal*)Sc);
                    }
                    star_rec(jlist.as<int>(), theta_c.as<double>(), recs_c, crs_c, ers_c, 0, 0);
                    render_launch<MODE_COST>(ldsg, dim3(gb_render), render_sh, M, recs_c.as<StarRec<T>>(),
                                             ers_c.as<EpsfRec<T>>(), crs_c.as<CornerRec<T>>(), Llive, g, ntiles, (const int2*)shares_l.as<int2>(),
                                             umask.as<uint8_t>(), dw.as<T>(), w.as<T>(), data.as<T>(), wres_cand,
                                             nullptr, part.as<double>(), Sc);
                    finish(k_finish<FIN_COST_CAND>, gb_render, 0, hS);   // and the scalars to the host
                };
                if (use_graphs) graphs.run(st, wres_cur, trial_body);
                else trial_body();
                CP_HIP(hipStreamSynchronize(st));
                if (trial == 1) {
                    if (have_live_count) {
                        n_live_now = hS->n_live;
                        if (n_live_now == 0) {   // every star froze at the last accepted step
                            none_live = true;
                            break;
                        }
                    }
                    cost = hS->cost;
                    const unsigned long long bits = hS->gmax_bits;
                    std::memcpy(&gnorm, &bits, 8);
                    if (gnorm <= o.g_tol) {
                        g_stop = true;
                        break;
                    }
                }
                const double cost_c = hS->cost_cand;
                out->n_trials_total += 1;
                out->n_inner_total += hS->iter;
                const double dcost = cost - cost_c;
                f_conv = (cost_c < cost) && (dcost / cost <= o.f_tol);
                if (out->trace && out->trace_len < out->trace_cap) {
                    double* t = out->trace + 8 * (size_t)out->trace_len++;
                    t[0] = iter;
                    t[1] = trial;
                    t[2] = cost;
                    t[3] = cost_c;
                    t[4] = lam;
                    t[5] = n_live_now;
                    t[6] = hS->iter;
                    t[7] = gnorm;
                }
                if (cost_c < cost) {
                    accepted = true;
                    cost = cost_c;
                    lam = std::max(lam / o.lambda_down, o.lambda_min);
                    break;
                }
                lam = std::min(lam * o.lambda_up, o.lambda_max);
            }
            if (none_live) {
                n_run = iter - 1;
                break;
            }
            if (g_stop) {
                launch(k_freeze_all, dim3(gstar), dim3(BLOCK), 0, st, na, iter - 1, live.as<uint8_t>(),
                       frozen_at.as<int>(), conv.as<uint8_t>());
                n_run = iter - 1;
                break;
            }
            n_run = iter;
            if (f_conv) {
                // theta <- theta + delta for the accepted step, then freeze everyone
                launch_n("k_accept", k_accept, dim3(gstar), dim3(BLOCK), 0, st, na, p, iter, -1.0, theta.as<double>(),
                       (const double*)theta_c.as<double>(), D.as<double>(), (const double*)colnorm.as<double>(),
                       (const double*)dsc.as<double>(), (const uint8_t*)gloc.as<uint8_t>(), live.as<uint8_t>(),
                       frozen_at.as<int>(), conv.as<uint8_t>(), part.as<double>());
                launch(k_freeze_all, dim3(gstar), dim3(BLOCK), 0, st, na, iter, live.as<uint8_t>(),
                       frozen_at.as<int>(), conv.as<uint8_t>());
                break;
            }
            if (accepted) {
                launch_n("k_accept", k_accept, dim3(gstar), dim3(BLOCK), 0, st, na, p, iter, o.x_tol, theta.as<double>(),
                       (const double*)theta_c.as<double>(), D.as<double>(), (const double*)colnorm.as<double>(),
                       (const double*)dsc.as<double>(), (const uint8_t*)gloc.as<uint8_t>(), live.as<uint8_t>(),
                       frozen_at.as<int>(), conv.as<uint8_t>(), part.as<double>());
                finish(k_finish<FIN_ACCEPT>, (int)gstar, 0);
                // the candidate's residual and records are the new point's (the
                // captured trial graphs follow, keyed by the residual buffer)
                std::swap(wres_cur, wres_cand);
                std::swap(recs, recs_c);
                std::swap(crs, crs_c);
                std::swap(ers, ers_c);
                have_live_count = true;
                // the stars frozen at this step (in the live list of the last
                // compaction) leave the data: records, tile marks and the freeze
                // render, each doing nothing when none froze
                star_rec(jlist.as<int>(), theta.as<double>(), recs, crs, ers, iter, 0);
                launch_n("k_mark_frozen", k_mark_frozen, dim3(gstar), dim3(BLOCK), 0, st, (const int*)jlist.as<int>(), (const Scal*)Sc,
                       (const uint8_t*)live.as<uint8_t>(), (const int*)frozen_at.as<int>(), iter,
                       (const int*)ay.as<int>(), (const int*)ax.as<int>(), (const int*)mR.as<int>(), g,
                       tmark.as<int>());
                render_launch<MODE_FREEZE>(ldsg, dim3(gb_render), render_sh, M, recs.as<StarRec<T>>(),
                                           ers.as<EpsfRec<T>>(), crs.as<CornerRec<T>>(), Llive, g, ntiles,
                                           (const int2*)shares_l.as<int2>(), umask.as<uint8_t>(),
                                           dw.as<T>(), w.as<T>(), data.as<T>(), wres_cur, nullptr,
                                           part.as<double>(), Sc, (const int*)tmark.as<int>(), iter);
            }
        }

        // --- epilogue: errors (:1302-1329), final render and chisq (:1344-1397)
        const int cov_mode = o.covariance == CP_COV_DEFAULT ? (d_iv ? CP_COV_KNOWN : CP_COV_REWEIGHTED) : o.covariance;
        star_rec(nullptr, theta.as<double>(), recs, crs, ers, 0, 1);
        fill(true);
        DevBuf errs(ctx, nb8);
        CP_HIP(hipMemsetAsync(errs.get(), 0, nb8, st));
        if (cov_mode != CP_COV_NONE) {
#define CP_ERR(PP)                                                                                             \
    launch_n("k_errors", k_errors<T, PP, VT>, dim3(cdiv(na, BLOCK / ERR_GL)), dim3(BLOCK), 0, st, na, R, S, (const int*)ay.as<int>(), \
           (const int*)ax.as<int>(), (const T*)w.as<T>(), ny, nx, (const VT*)V.as<VT>(),                        \
           (const double*)colnorm.as<double>(), cov_mode, cost / dof, errs.as<double>())
            if (p == 1) CP_ERR(1);
            else if (p == 2) CP_ERR(2);
            else CP_ERR(3);
#undef CP_ERR
        }
        T* resid = ubuf.as<T>();
        render_launch<MODE_FINAL>(ldsg, dim3(gb_render), render_sh, M, recs.as<StarRec<T>>(), ers.as<EpsfRec<T>>(),
                                  crs.as<CornerRec<T>>(),
                                  Lmod, g, ntiles, (const int2*)shares_m.as<int2>(), umask.as<uint8_t>(), dw.as<T>(), w.as<T>(),
                                  data.as<T>(), wres_cur, resid, part.as<double>(), Sc);
        // --- per-catalogue-star results, assembled on the device (:1289-1300,
        // :1331-1337, :1414-1430)
        const size_t N = (size_t)n;
        DevBuf slot(ctx, N * 4);
        CP_HIP(hipMemsetAsync(slot.get(), 0xff, N * 4, st));   // -1: not active
        launch(k_slot_of, dim3(cdiv(na, BLOCK)), dim3(BLOCK), 0, st, (const int*)act.as<int>(), na, slot.as<int>());
        DevBuf res(ctx, N * (OUT_F64 * 8 + 4 + 1 + 1));
        double* cols = res.as<double>();
        int* d_niter = reinterpret_cast<int*>(cols + OUT_F64 * N);
        uint8_t* d_conv = reinterpret_cast<uint8_t*>(d_niter + N);
        uint8_t* d_valid = d_conv + N;
        launch_n("k_results", k_results, dim3(cdiv(n, BLOCK)), dim3(BLOCK), 0, st, n, (const int*)slot.as<int>(), d_y0, d_x0, d_f0,
               (const double*)theta.as<double>(), (const double*)errs.as<double>(), cov_mode != CP_COV_NONE ? 1 : 0,
               (const uint8_t*)live.as<uint8_t>(), (const uint8_t*)conv.as<uint8_t>(),
               (const int*)frozen_at.as<int>(), n_run, fm, o.fixed_bkg, cols, d_niter, d_conv, d_valid);

        // --- per-star diagnostics over the global residual (:1344-1397,
        // psf_photometry_diagnostics.jl:154-231); the spread_model disk is
        // sized from catalogue star 1's reported position unless given
        double* host_cols[OUT_F64] = {out->y,        out->x,         out->flux,   out->bkg,      out->y_err,
                                      out->x_err,    out->flux_err,  out->bkg_err, out->chisq,   out->qfit,
                                      out->qfit_expected, out->qfit_z, out->crowding, out->spread_model,
                                      out->spread_model_err};
        bool any_diag = false;
        for (int k = OUT_CHISQ; k < OUT_F64; ++k) any_diag = any_diag || host_cols[k];
        DevBuf dK;
        if (any_diag) {
            double sfw = NAN;
            if (out->spread_model || out->spread_model_err) {
                if (o.spread_model_fwhm > 0) {
                    sfw = o.spread_model_fwhm;
                } else {
                    double yx[2];
                    CP_HIP(hipMemcpyAsync(&yx[0], cols + OUT_Y * N, 8, hipMemcpyDeviceToHost, st));
                    CP_HIP(hipMemcpyAsync(&yx[1], cols + OUT_X * N, 8, hipMemcpyDeviceToHost, st));
                    CP_HIP(hipStreamSynchronize(st));
                    sfw = spread_fwhm_host(psfd_, yx[0], yx[1]);
                }
            }
            const int kh = (int)std::ceil(fit_rad);
            std::vector<double> hK;
            if (std::isfinite(sfw) && sfw > 0) {
                hK.resize((size_t)(2 * kh + 1) * (2 * kh + 1));
                spread_kernel_host(sfw, kh, hK.data());
                dK.alloc(ctx, hK.size() * 8);
                CP_HIP(hipMemcpyAsync(dK.get(), hK.data(), hK.size() * 8, hipMemcpyHostToDevice, st));
            }
            const int dmaxb = (2 * (int)std::ceil(fit_rad) + 2) * (2 * (int)std::ceil(fit_rad) + 2);
            launch_n("k_diag", k_diag<T, KIND>, dim3(cdiv(na, 64 / DIAG_GL)), dim3(64),
                   (size_t)(64 / DIAG_GL) * dmaxb * sizeof(T), st, na, (const double*)theta.as<double>(), fm,
                   fit_rad, M, inv_norm_g64, moffat_amp64, DP.G, (const T*)resid, (const T*)data.as<T>(),
                   (const T*)w.as<T>(), d_iv ? 1 : 0, ny, nx, (const double*)(dK.get() ? dK.as<double>() : nullptr),
                   kh, (const int*)act.as<int>(), n, dmaxb, cols);
            if (!hK.empty()) CP_HIP(hipStreamSynchronize(st));   // hK is pageable and local
        }

        // --- read-back: every requested column and the residual image
        std::vector<D2HSeg> segs;
        for (int k = 0; k < OUT_F64; ++k)
            if (host_cols[k]) segs.push_back({host_cols[k], cols + k * N, N * 8});
        if (out->n_iter) segs.push_back({out->n_iter, d_niter, N * 4});
        if (out->converged) segs.push_back({out->converged, d_conv, N});
        if (out->valid) segs.push_back({out->valid, d_valid, N});
        DevBuf rf;
        if (out->residual) {
            const void* rsrc = resid;
            if (o.layout == CP_LAYOUT_ROWMAJOR) {
                // column-major ny x nx is row-major nx x ny: transpose to ny x nx
                rf.alloc(ctx, Pn * 4);
                launch(k_transpose<T, float>, dim3(cdiv(ny, 32), cdiv(nx, 32)), dim3(32, 8), 0, st,
                       (const T*)resid, nx, ny, rf.as<float>());
                rsrc = rf.get();
            } else if (sizeof(T) != 4) {
                rf.alloc(ctx, Pn * 4);
                launch(k_to_float<T>, dim3(std::min<int64_t>(cdiv(Pn, BLOCK), n_cu * 16)), dim3(BLOCK), 0, st,
                       (const T*)resid, rf.as<float>(), Pn);
                rsrc = rf.get();
            }
            segs.push_back({out->residual, rsrc, (size_t)Pn * 4});
        }
        staged_d2h(ctx, st, segs);
        out->n_passes = n_run;
        out->n_active = na;
        out->n_union = U;
        out->cost_final = cost;
    }
};

void validate_sim(const cp_sim_opts& o, double fit_rad) {
    CP_ARG(o.max_iter > 0, "max_iter must be positive");
    CP_ARG(o.inner_iterations > 0, "inner_iterations must be positive");
    CP_ARG(o.linear_tol > 0, "linear_tol must be positive");
    CP_ARG(o.max_trials > 0, "max_trials must be positive");
    double mrm = o.model_rad_max > 0 ? o.model_rad_max : 10 * fit_rad;
    CP_ARG(mrm >= fit_rad, "model_rad_max must be >= fit_rad");
    CP_ARG(o.model_rad_nsigma > 0, "model_rad_nsigma must be positive");
    CP_ARG(o.solver == CP_SOLVER_LSQR || o.solver == CP_SOLVER_LSMR || o.solver == CP_SOLVER_CGLS,
           "solver must be CP_SOLVER_LSQR, CP_SOLVER_LSMR or CP_SOLVER_CGLS");
    CP_ARG(fit_rad > 0 && std::isfinite(fit_rad), "fit_rad must be positive");
    CP_ARG((o.free_mask & 7) != 0, "all model parameters are fixed; nothing to fit");
    CP_ARG(o.precision == CP_PREC_MIXED || o.precision == CP_PREC_FP64, "unknown precision mode");
}

void fit_sim_device(cp_ctx* ctx, const cp_psf_desc* psfd, const float* d_img, const float* d_iv, int ny, int nx,
                    const double* d_y0, const double* d_x0, const double* d_f0, int n, double fit_rad,
                    const cp_sim_opts* po, cp_sim_result* out) {
    CP_ARG(po && out, "NULL options or result");
    CP_ARG(ny > 0 && nx > 0, "image must be non-empty");
    CP_ARG(n >= 0, "negative star count");
    const cp_sim_opts& o = *po;
    validate_sim(o, fit_rad);
    CP_ARG(o.layout == CP_LAYOUT_COLMAJOR || o.layout == CP_LAYOUT_ROWMAJOR, "layout must be CP_LAYOUT_*");
    HostPsf P = make_host_psf(psfd);
    CP_HIP(hipSetDevice(ctx->device));
    DevPsf DP;
    upload_psf(ctx, psfd, P, DP);
    DevBuf t_img, t_iv;
    if (o.layout == CP_LAYOUT_ROWMAJOR) {
        // C-order inputs: transpose into the column-major working layout
        const size_t Pn = (size_t)ny * nx;
        t_img.alloc(ctx, Pn * 4);
        launch(k_transpose<float, float>, dim3(cdiv(nx, 32), cdiv(ny, 32)), dim3(32, 8), 0, ctx->stream, d_img, ny, nx,
               t_img.as<float>());
        d_img = t_img.as<float>();
        if (d_iv) {
            t_iv.alloc(ctx, Pn * 4);
            launch(k_transpose<float, float>, dim3(cdiv(nx, 32), cdiv(ny, 32)), dim3(32, 8), 0, ctx->stream, d_iv,
                   ny, nx, t_iv.as<float>());
            d_iv = t_iv.as<float>();
        }
    }
    std::vector<double> hy(n), hx(n), hf(n);
    if (n > 0) {
        CP_HIP(hipMemcpyAsync(hy.data(), d_y0, (size_t)n * 8, hipMemcpyDeviceToHost, ctx->stream));
        CP_HIP(hipMemcpyAsync(hx.data(), d_x0, (size_t)n * 8, hipMemcpyDeviceToHost, ctx->stream));
        CP_HIP(hipMemcpyAsync(hf.data(), d_f0, (size_t)n * 8, hipMemcpyDeviceToHost, ctx->stream));
        CP_HIP(hipStreamSynchronize(ctx->stream));
    }
#define CP_RUN(TT, KK) Sim<TT, KK>(ctx, P, DP, psfd).run(d_img, d_iv, ny, nx, d_y0, d_x0, d_f0, n, fit_rad, o, out, hy, hx, hf)
    const bool f64 = o.precision == CP_PREC_FP64;
    switch (P.kind) {
        case CP_PSF_CIRCULAR_GAUSSIAN:
            if (f64) CP_RUN(double, KIND_GAUSS); else CP_RUN(float, KIND_GAUSS);
            break;
        case CP_PSF_CIRCULAR_MOFFAT:
            if (f64) CP_RUN(double, KIND_MOFFAT); else CP_RUN(float, KIND_MOFFAT);
            break;
        default:
            if (f64) CP_RUN(double, KIND_EPSF); else CP_RUN(float, KIND_EPSF);
            break;
    }
#undef CP_RUN
    CP_HIP(hipStreamSynchronize(ctx->stream));
}

}  // namespace cp

using namespace cp;

extern "C" {

void cp_sim_opts_default(cp_sim_opts* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->solver = CP_SOLVER_LSQR;
    o->inner_iterations = 10;
    o->linear_tol = 1e-6;
    o->model_rad = 0;
    o->model_rad_max = 0;
    o->model_rad_nsigma = 1.0;
    o->max_step = 0.25;
    o->max_trials = 8;
    o->max_iter = 40;
    o->x_tol = 1e-5;
    o->f_tol = 1e-4;
    o->g_tol = 1e-4;
    o->lambda_init = 1e-4;
    o->lambda_up = 10.0;
    o->lambda_down = 10.0;
    o->lambda_min = 1e-12;
    o->lambda_max = 1e12;
    o->covariance = CP_COV_DEFAULT;
    o->free_mask = CP_FREE_Y | CP_FREE_X | CP_FREE_FLUX;
    o->precision = CP_PREC_MIXED;
    o->tile = 0;
}

int cp_fit_simultaneous_device(cp_ctx* ctx, const cp_psf_desc* psf, const float* d_image, const float* d_inv_var,
                               int32_t ny, int32_t nx, const double* d_y0, const double* d_x0,
                               const double* d_flux0, int32_t n_stars, double fit_rad, const cp_sim_opts* opts,
                               cp_sim_result* out) {
    if (!ctx) return CP_ESTATE;
    return guarded(ctx, [&] {
        fit_sim_device(ctx, psf, d_image, d_inv_var, ny, nx, d_y0, d_x0, d_flux0, n_stars, fit_rad, opts, out);
    });
}

int cp_fit_simultaneous(cp_ctx* ctx, const cp_psf_desc* psf, const float* image, const float* inv_var, int32_t ny,
                        int32_t nx, const double* y0, const double* x0, const double* flux0, int32_t n_stars,
                        double fit_rad, const cp_sim_opts* opts, cp_sim_result* out) {
    if (!ctx) return CP_ESTATE;
    return guarded(ctx, [&] {
        CP_ARG(ny > 0 && nx > 0, "image must be non-empty");
        CP_ARG(image != nullptr, "image is NULL");
        CP_HIP(hipSetDevice(ctx->device));
        const size_t P = (size_t)ny * nx;
        DevBuf di(ctx, P * 4), dv, dy(ctx, (size_t)std::max(n_stars, 1) * 8), dx(ctx, (size_t)std::max(n_stars, 1) * 8),
            df(ctx, (size_t)std::max(n_stars, 1) * 8);
        CP_HIP(hipMemcpyAsync(di.get(), image, P * 4, hipMemcpyHostToDevice, ctx->stream));
        if (inv_var) {
            dv.alloc(ctx, P * 4);
            CP_HIP(hipMemcpyAsync(dv.get(), inv_var, P * 4, hipMemcpyHostToDevice, ctx->stream));
        }
        if (n_stars > 0) {
            CP_HIP(hipMemcpyAsync(dy.get(), y0, (size_t)n_stars * 8, hipMemcpyHostToDevice, ctx->stream));
            CP_HIP(hipMemcpyAsync(dx.get(), x0, (size_t)n_stars * 8, hipMemcpyHostToDevice, ctx->stream));
            CP_HIP(hipMemcpyAsync(df.get(), flux0, (size_t)n_stars * 8, hipMemcpyHostToDevice, ctx->stream));
        }
        fit_sim_device(ctx, psf, di.as<float>(), inv_var ? dv.as<float>() : nullptr, ny, nx, dy.as<double>(),
                       dx.as<double>(), df.as<double>(), n_stars, fit_rad, opts, out);
    });
}

int cp_build_stamps(cp_ctx* ctx, const float* w_in, int32_t ny, int32_t nx, const double* y0, const double* x0,
                    const double* flux0, int32_t n_stars, double fit_rad, int32_t* anchor_y, int32_t* anchor_x,
                    int32_t* active, int32_t* n_active, int32_t* pixels, uint8_t* union_mask) {
    if (!ctx) return CP_ESTATE;
    return guarded(ctx, [&] {
        CP_ARG(ny > 0 && nx > 0 && n_stars >= 0, "bad sizes");
        CP_ARG(fit_rad > 0, "fit_rad must be positive");
        CP_HIP(hipSetDevice(ctx->device));
        hipStream_t st = ctx->stream;
        const int64_t P = (int64_t)ny * nx;
        const int R = (int)std::ceil(fit_rad), S = 2 * R + 1, S2 = S * S;
        DevBuf w(ctx, P * 4);
        if (w_in) {
            CP_HIP(hipMemcpyAsync(w.get(), w_in, P * 4, hipMemcpyHostToDevice, st));
        } else {
            std::vector<float> ones(P, 1.0f);
            CP_HIP(hipMemcpyAsync(w.get(), ones.data(), P * 4, hipMemcpyHostToDevice, st));
            CP_HIP(hipStreamSynchronize(st));
        }
        const int n = std::max(n_stars, 1);
        DevBuf dy(ctx, (size_t)n * 8), dx(ctx, (size_t)n * 8), df(ctx, (size_t)n * 8);
        if (n_stars > 0) {
            CP_HIP(hipMemcpyAsync(dy.get(), y0, (size_t)n_stars * 8, hipMemcpyHostToDevice, st));
            CP_HIP(hipMemcpyAsync(dx.get(), x0, (size_t)n_stars * 8, hipMemcpyHostToDevice, st));
            CP_HIP(hipMemcpyAsync(df.get(), flux0, (size_t)n_stars * 8, hipMemcpyHostToDevice, st));
        }
        DevBuf ay(ctx, (size_t)n * 4), ax(ctx, (size_t)n * 4), ok(ctx, n);
        *n_active = 0;
        if (n_stars > 0) {
            launch(k_anchor<float>, dim3(cdiv(n_stars, BLOCK)), dim3(BLOCK), 0, st, dy.as<double>(), dx.as<double>(),
                   df.as<double>(), n_stars, (const float*)w.as<float>(), ny, nx, R, ay.as<int>(), ax.as<int>(),
                   ok.as<uint8_t>());
            Cub cub{ctx, DevBuf()};
            DevBuf act;
            int coord_bits = bit_length((int64_t)std::max(ny, nx) + R + 1);
            int na = build_active(ctx, cub, ay.as<int>(), ax.as<int>(), ok.as<uint8_t>(), n_stars, coord_bits, act);
            *n_active = na;
            CP_HIP(hipMemcpyAsync(anchor_y, ay.get(), (size_t)n_stars * 4, hipMemcpyDeviceToHost, st));
            CP_HIP(hipMemcpyAsync(anchor_x, ax.get(), (size_t)n_stars * 4, hipMemcpyDeviceToHost, st));
            if (na > 0) {
                CP_HIP(hipMemcpyAsync(active, act.get(), (size_t)na * 4, hipMemcpyDeviceToHost, st));
                DevBuf aay(ctx, (size_t)na * 4), aax(ctx, (size_t)na * 4), amr(ctx, (size_t)na * 4),
                    th(ctx, (size_t)na * 8), lv(ctx, na), fz(ctx, (size_t)na * 4), cv(ctx, na);
                launch(k_star_setup, dim3(cdiv(na, BLOCK)), dim3(BLOCK), 0, st, (const int*)act.as<int>(), na,
                       (const int*)ay.as<int>(), (const int*)ax.as<int>(), (const double*)dy.as<double>(),
                       (const double*)dx.as<double>(), (const double*)df.as<double>(), 1, 3, 0, 0, (const int*)nullptr,
                       R, aay.as<int>(), aax.as<int>(), amr.as<int>(), th.as<double>(), lv.as<uint8_t>(),
                       fz.as<int>(), cv.as<uint8_t>());
                if (pixels) {
                    DevBuf pix(ctx, (size_t)na * S2 * 4);
                    launch(k_pixels<float>, dim3(cdiv((int64_t)na * S2, BLOCK)), dim3(BLOCK), 0, st, na,
                           (const int*)aay.as<int>(), (const int*)aax.as<int>(), R, S, (const float*)w.as<float>(),
                           ny, nx, pix.as<int>());
                    CP_HIP(hipMemcpyAsync(pixels, pix.get(), (size_t)na * S2 * 4, hipMemcpyDeviceToHost, st));
                    CP_HIP(hipStreamSynchronize(st));
                }
                if (union_mask) {
                    DevBuf um(ctx, P);
                    CP_HIP(hipMemsetAsync(um.get(), 0, P, st));
                    launch(k_union<float>, dim3(cdiv((int64_t)na * S2, BLOCK)), dim3(BLOCK), 0, st, na,
                           (const int*)aay.as<int>(), (const int*)aax.as<int>(), R, S, (const float*)w.as<float>(), ny,
                           nx, um.as<uint8_t>());
                    CP_HIP(hipMemcpyAsync(union_mask, um.get(), P, hipMemcpyDeviceToHost, st));
                    CP_HIP(hipStreamSynchronize(st));
                }
            }
            CP_HIP(hipStreamSynchronize(st));
        }
        if (union_mask && *n_active == 0) std::memset(union_mask, 0, P);
    });
}

}  // extern "C"
