"""Launch shapes (cp_tuning) on the GPU: range checks, the batched per-star
LM unchanged bit for bit under another grid, the simultaneous fit under
other shapes and work-tile sizes (render tile side, lanes per normal-matrix
block and per star of the fill) still on the FP64 oracle, folded Golub-Kahan
steps bit for bit equal to unfolded ones, and the autotuner on a real call."""
import numpy as np
import pytest

from _support import load_pkg

pkg = load_pkg()
sim = pkg.simulation

pytestmark = pytest.mark.gpu


@pytest.fixture()
def ctx():
    c = pkg.Context(0)
    yield c
    c.close()


def test_tuning_round_trip_and_range(ctx):
    assert set(ctx.get_tuning().values()) == {0}
    ctx.set_tuning(hbuild_blocks_per_cu=4, lm_blocks_per_cu=2)
    t = ctx.get_tuning()
    assert t["hbuild_blocks_per_cu"] == 4 and t["lm_blocks_per_cu"] == 2 and t["cg_blocks_per_cu"] == 0
    with pytest.raises(pkg.ArgumentError, match="blocks per CU"):
        ctx.set_tuning(render_blocks_per_cu=65)
    for field, bad in (("render_tile", 48), ("hbuild_lanes", 8), ("fill_lanes", 32), ("lm_work_tile", 65),
                       ("gk_fold_blocks", -1)):
        with pytest.raises(pkg.ArgumentError, match=field):
            ctx.set_tuning(**{field: bad})
    with pytest.raises(pkg.ArgumentError, match="unknown tuning"):
        ctx.set_tuning(tile=3)
    assert ctx.get_tuning() == t


def test_batched_lm_identical_under_other_grids(ctx):
    c = sim.config_cutouts(seed=3, n=3000)
    model = pkg.CircularMoffatPSF(alpha=1.0, beta=c["beta"])
    cut = [c["pix"][o:o + s * s].reshape(s, s).T for o, s in zip(c["offsets"], c["ny"])]
    ref = pkg.fit_star_batch(model, cut, c["theta0"], fixed=("alpha", "beta"), ctx=ctx)
    for bpc, tile in ((1, 1), (2, 4), (16, 2), (8, 64)):
        ctx.set_tuning(lm_blocks_per_cu=bpc, lm_work_tile=tile)
        got = pkg.fit_star_batch(model, cut, c["theta0"], fixed=("alpha", "beta"), ctx=ctx)
        np.testing.assert_array_equal(got[0], ref[0])
        np.testing.assert_array_equal(got[1], ref[1])


def test_simultaneous_under_other_shapes_matches_oracle(ctx):
    """The Gaussian-field parity case (256x256, 200 stars) and a crowded ePSF
    crop under launch shapes far from the defaults: the same bars as with the
    default shape."""
    from test_gpu_parity import FLUX_RTOL, POS_ATOL, _compare, _gpu_run, _oracle_run
    cfg = sim.config_gauss_small(seed=1)
    cfg["image"] = cfg["image"].astype(np.float32).astype(np.float64)
    ro = _oracle_run(cfg)
    n = int(round(2e5 * (384 * 384) / 4096 ** 2))
    crowd = sim.config_epsf_crowded(seed=5, shape=(384, 384), n=n, max_iter=10)
    ref = _gpu_run(ctx, crowd, "mixed")
    ref64 = _gpu_run(ctx, crowd, "fp64")
    for shape in ({"hbuild_blocks_per_cu": 1, "cg_blocks_per_cu": 16, "render_blocks_per_cu": 1,
                   "render_tile": 32, "hbuild_lanes": 64, "fill_lanes": 64},
                  {"hbuild_blocks_per_cu": 32, "cg_blocks_per_cu": 2, "render_blocks_per_cu": 8,
                   "render_tile": 64, "hbuild_lanes": 16, "fill_lanes": 16}):
        ctx.set_tuning(**shape)
        _compare(_gpu_run(ctx, cfg, "fp64"), ro, 1e-7, 1e-7)
        _compare(_gpu_run(ctx, cfg, "mixed"), ro, FLUX_RTOL, POS_ATOL)
        # crowded field: against the default shape's own run (in float64 the
        # shapes regroup only rounding-level partial sums: every star agrees)
        _compare(_gpu_run(ctx, crowd, "mixed"), ref, FLUX_RTOL, POS_ATOL, frac=0.98)
        _compare(_gpu_run(ctx, crowd, "fp64"), ref64, 1e-8, 1e-8)


@pytest.mark.parametrize("solver", ["lsqr", "lsmr"])
def test_folded_golub_kahan_steps_identical(ctx, solver):
    """Golub-Kahan steps with each step's reduction folded into the next
    launch give the results of one reduction launch per step, bit for bit:
    a 512x512 tile (a grid of several blocks, folded by default) with
    folding off, and the 4096x4096-density crop with folding forced on."""
    t = sim.serving_tile(3, 512, "gauss")
    kw = dict(fixed=t["fixed"], inv_var=t["inv_var"], model_rad=t["model_rad"], max_iter=t["max_iter"],
              solver=solver, ctx=ctx)
    n = int(round(2e5 * (384 * 384) / 4096 ** 2))
    crowd = sim.config_epsf_crowded(seed=5, shape=(384, 384), n=n, max_iter=10)
    ref = pkg.fit_all_stars_simultaneous(t["image"], t["psf"], t["sources"], t["fit_rad"], **kw)
    refc = _gpu_run_solver(ctx, crowd, solver)
    for fold in (1, 1 << 20):
        ctx.set_tuning(gk_fold_blocks=fold)
        got = pkg.fit_all_stars_simultaneous(t["image"], t["psf"], t["sources"], t["fit_rad"], **kw)
        gotc = _gpu_run_solver(ctx, crowd, solver)
        for a, b in ((got, ref), (gotc, refc)):
            for f in ("y", "x", "flux", "flux_err", "chisq", "n_iter", "valid", "converged"):
                np.testing.assert_array_equal(getattr(a, f), getattr(b, f), err_msg=f"{f} fold {fold}")
            assert a.n_passes == b.n_passes
    # the shortest solves: one step (the closing reduction reads the first
    # step's state copy) and two
    for inner in (1, 2):
        ctx.set_tuning(gk_fold_blocks=0)
        ref = pkg.fit_all_stars_simultaneous(t["image"], t["psf"], t["sources"], t["fit_rad"],
                                             inner_iterations=inner, **kw)
        ctx.set_tuning(gk_fold_blocks=1)
        got = pkg.fit_all_stars_simultaneous(t["image"], t["psf"], t["sources"], t["fit_rad"],
                                             inner_iterations=inner, **kw)
        for f in ("y", "x", "flux", "flux_err", "n_iter"):
            np.testing.assert_array_equal(getattr(got, f), getattr(ref, f), err_msg=f"{f} inner {inner}")


def _gpu_run_solver(ctx, cfg, solver):
    from test_gpu_parity import _gpu_run
    return _gpu_run(ctx, cfg, "mixed", solver=solver)


def test_autotune_on_a_real_call(ctx):
    c = sim.config_cutouts(seed=5, n=2000)
    model = pkg.CircularMoffatPSF(alpha=1.0, beta=c["beta"])
    cut = [c["pix"][o:o + s * s].reshape(s, s).T for o, s in zip(c["offsets"], c["ny"])]
    ref = pkg.fit_star_batch(model, cut, c["theta0"], fixed=("alpha", "beta"), ctx=ctx)

    def run():
        pkg.fit_star_batch(model, cut, c["theta0"], fixed=("alpha", "beta"), ctx=ctx)

    shape, t = pkg.autotune.tune(ctx, run, space=pkg.autotune.LM_SPACE, repeats=1)
    assert shape["lm_blocks_per_cu"] in pkg.autotune.LM_SPACE["lm_blocks_per_cu"] and t > 0
    assert ctx.get_tuning()["lm_blocks_per_cu"] == shape["lm_blocks_per_cu"]
    got = pkg.fit_star_batch(model, cut, c["theta0"], fixed=("alpha", "beta"), ctx=ctx)
    np.testing.assert_array_equal(got[0], ref[0])
