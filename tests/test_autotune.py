"""The launch-shape autotuner's search and cache (CPU; a stand-in context
records the shapes, and the timed call sleeps for a cost that depends on the
shape, so the search has a known optimum)."""
import json
import time

from _support import load_pkg

pkg = load_pkg()
at = pkg.autotune


class FakeCtx:
    def __init__(self):
        self.t = {k: 0 for k in ("hbuild_blocks_per_cu", "cg_blocks_per_cu", "render_blocks_per_cu",
                                 "lm_blocks_per_cu", "render_tile", "hbuild_lanes", "fill_lanes",
                                 "lm_work_tile", "gk_fold_blocks")}
        self.history = []

    def get_tuning(self):
        return dict(self.t)

    def set_tuning(self, **kw):
        self.t.update(kw)
        self.history.append(dict(self.t))


def _run_for(ctx, best):
    def run():
        cost = sum(abs(ctx.t[k] - v) for k, v in best.items())
        time.sleep(0.002 + 0.004 * cost / 16)
    return run


def test_coordinate_descent_finds_the_known_optimum():
    ctx = FakeCtx()
    best = {"hbuild_blocks_per_cu": 16, "render_blocks_per_cu": 2}
    shape, t = at.tune(ctx, _run_for(ctx, best), repeats=2)
    for k, v in best.items():
        assert shape[k] == v, (k, shape)
    assert ctx.get_tuning()["hbuild_blocks_per_cu"] == 16   # left set on the context
    assert t < 0.01


def test_work_tile_sizes_are_searched():
    ctx = FakeCtx()
    best = {"render_tile": 32, "hbuild_lanes": 16, "fill_lanes": 64}
    shape, _ = at.tune(ctx, _run_for(ctx, best), repeats=2)
    for k, v in best.items():
        assert shape[k] == v, (k, shape)
    # 0 (the stamp-size and image-size rules) is where the H-assembly lanes and the render tile start
    assert ctx.history[0]["hbuild_lanes"] == 0 and ctx.history[0]["render_tile"] == 0


def test_defaults_are_the_starting_point_and_kept_without_gain():
    ctx = FakeCtx()
    best = dict(at.DEFAULTS)
    shape, _ = at.tune(ctx, _run_for(ctx, {k: best[k] for k in at.LM_SPACE}), space=at.LM_SPACE, repeats=2)
    assert shape["lm_blocks_per_cu"] == at.DEFAULTS["lm_blocks_per_cu"]
    assert ctx.history[0]["lm_blocks_per_cu"] == at.DEFAULTS["lm_blocks_per_cu"]


def test_cache_round_trip(tmp_path):
    path = str(tmp_path / "tuning.json")
    ctx = FakeCtx()
    best = {"hbuild_blocks_per_cu": 8, "render_blocks_per_cu": 8}
    shape = at.cached(ctx, "crowded/4096", _run_for(ctx, best), path)
    assert {k: shape[k] for k in best} == best
    assert json.load(open(path))["crowded/4096"] == shape
    ctx2 = FakeCtx()
    calls = []
    shape2 = at.cached(ctx2, "crowded/4096", lambda: calls.append(1), path)
    assert shape2 == shape and calls == [] and ctx2.get_tuning() == shape


def test_failing_call_restores_the_shape():
    ctx = FakeCtx()
    ctx.set_tuning(hbuild_blocks_per_cu=16)
    before = ctx.get_tuning()
    n = []

    def run():
        n.append(1)
        if len(n) > 3:
            raise RuntimeError("device error")

    try:
        at.tune(ctx, run)
    except RuntimeError:
        pass
    else:
        raise AssertionError("the call's error must propagate")
    assert ctx.get_tuning() == before
