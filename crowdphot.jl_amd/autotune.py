"""Launch-shape autotuner for the grid-stride kernels.

The tunable shapes are the cp_tuning fields of a context: blocks per compute
unit of the normal-matrix assembly, CG's per-star update, the tile-gather
renders and the batched per-star LM, and the simultaneous fit's work-tile
sizes (the render tile side, lanes per normal-matrix block, lanes per star of
the fill) and the batched LM's work tile (cutouts per work-counter increment).  `tune` times complete calls of a workload -- whatever
`run` does, typically one device-resident fit -- over candidate values, one
field at a time (coordinate descent from the current shape, best of `repeats`
wall-clock timings per candidate, the call synchronous), leaves the best
shape set on the context and returns it.  `cached` keeps the chosen shapes in
a small JSON file keyed by a workload name, so a later process applies them
without timing.

Shapes change no result of the batched per-star LM (every cutout's fit is
independent of the grid) or of the normal-matrix assembly's grid (one group
per block, whatever the grid).  CG's update shape and the render shape and
tile side regroup per-block partial sums (CG's norms, the costs), and the
lanes per block or per star regroup the sums of H and of J^T r: a
rounding-level difference that is fixed for a fixed shape.
"""
import json
import os
import time

DEFAULT_SPACE = {
    "hbuild_blocks_per_cu": (8, 16, 32),
    "render_blocks_per_cu": (2, 4, 8),
    "render_tile": (32, 64),
    "hbuild_lanes": (4, 16, 64),
    "fill_lanes": (16, 64),
}
LM_SPACE = {"lm_blocks_per_cu": (2, 4, 8, 16), "lm_work_tile": (1, 2, 4, 8)}
# the built-in defaults (what 0 stands for), as starting points; 0 for the
# H-assembly lanes, whose default follows the stamp size, and for the render
# tile, whose default follows the image size
DEFAULTS = {"hbuild_blocks_per_cu": 32, "cg_blocks_per_cu": 8, "render_blocks_per_cu": 2, "lm_blocks_per_cu": 8,
            "render_tile": 0, "hbuild_lanes": 0, "fill_lanes": 16, "lm_work_tile": 1, "gk_fold_blocks": 0}


def _time(run, repeats):
    best = float("inf")
    for _ in range(repeats):
        t = time.perf_counter()
        run()
        best = min(best, time.perf_counter() - t)
    return best


def tune(ctx, run, space=None, repeats=2, rounds=1, log=None):
    """Coordinate descent over `space` (field -> candidate values); `run()`
    performs one complete, synchronous call.  Returns (shape, seconds) with
    the shape set on ctx."""
    space = DEFAULT_SPACE if space is None else space
    before = ctx.get_tuning()
    shape = dict(before)
    for k in space:
        if shape.get(k, 0) == 0:
            shape[k] = DEFAULTS[k]
    try:
        ctx.set_tuning(**shape)
        run()   # warm-up: allocations, code loading
        best_t = _time(run, repeats)
        for _ in range(rounds):
            changed = False
            for field, cands in space.items():
                for v in cands:
                    if v == shape[field]:
                        continue
                    trial = dict(shape, **{field: v})
                    ctx.set_tuning(**trial)
                    t = _time(run, repeats)
                    if log:
                        log(f"[tune] {field}={v}: {t * 1e3:.2f} ms (best {best_t * 1e3:.2f} ms)")
                    if t < best_t:
                        best_t, shape, changed = t, trial, True
                ctx.set_tuning(**shape)
            if not changed:
                break
    except BaseException:
        ctx.set_tuning(**before)   # a failing call leaves the context as it was
        raise
    ctx.set_tuning(**shape)
    return shape, best_t


def cached(ctx, key, run, path, space=None, repeats=2, log=None):
    """The shape stored under `key` in the JSON file `path` (set on ctx), or a
    fresh `tune` whose result is stored there."""
    store = {}
    if os.path.exists(path):
        with open(path) as f:
            store = json.load(f)
    if key in store:
        ctx.set_tuning(**store[key])
        return store[key]
    shape, _ = tune(ctx, run, space=space, repeats=repeats, log=log)
    store[key] = shape
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(store, f, indent=1, sort_keys=True)
    os.replace(tmp, path)
    return shape
