#!/usr/bin/env python3
"""Headline benchmark: simultaneous PSF photometry of a crowded ePSF field on
one MI355X.

Workload (one "step" = one complete fit_all_stars_simultaneous call):
4096 x 4096 float32 image, 200,000 stars, ePSF 101 x 101 at 4x oversampling
(Moffat beta 2.5, FWHM 2.2 px), fit_rad 3 (7 x 7 stamps), model_rad 12,
10 outer Gauss-Newton iterations, LSQR inner solves (10 iterations), inputs
(image, inv_var, catalogue) resident in HBM before the timed region.

Prints one JSON line.  value = stars fitted per second (stars / wall time of
a fit).  Extra fields: roofline of the dominant kernel (HIP-event timing of
every launch during one instrumented fit), accuracy against the float64
device run, and the CPU baseline (the FP64 oracle on a bounded crop).

    python bench.py [--gpus 1] [--steps 10] [--warmup 2] [--no-cpu-baseline]
"""
import argparse
import ctypes as C
import importlib.util
import json
import math
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "crowdphot.jl_amd")
HBM_PEAK_GBS = 8000.0
# every per-star output of fit_all_stars_simultaneous's result
# (MultiPassPhotResult, src/photometry/psf_photometry_single.jl:99-122)
PER_STAR = ("y", "x", "flux", "bkg", "y_err", "x_err", "flux_err", "bkg_err", "chisq", "qfit", "qfit_expected",
            "qfit_z", "crowding", "spread_model", "spread_model_err")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def load_pkg():
    if "crowdphot_amd" in sys.modules:
        return sys.modules["crowdphot_amd"]
    spec = importlib.util.spec_from_file_location(
        "crowdphot_amd", os.path.join(PKG_DIR, "__init__.py"), submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["crowdphot_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def star_pairs(y0, x0, R):
    """Pairs of catalogue stars whose (2R+1)^2 stamps overlap (rounded anchors
    within 2R on both axes), each star with itself included: the rows of the
    star-space normal matrix."""
    ay, ax = np.rint(y0).astype(np.int64), np.rint(x0).astype(np.int64)
    cs = 2 * R
    cy, cx = (ay + R - 1) // cs, (ax + R - 1) // cs
    order = np.lexsort((cy, cx))
    key = cx[order] * (cy.max() + 3) + cy[order]
    total = 0
    for dx in (-1, 0, 1):
        for dy in (-1, 0, 1):
            k2 = (cx + dx) * (cy.max() + 3) + (cy + dy)
            lo, hi = np.searchsorted(key, k2, "left"), np.searchsorted(key, k2, "right")
            c = hi - lo
            i = np.repeat(np.arange(len(ay)), c)
            start = np.repeat(lo - np.concatenate(([0], np.cumsum(c)[:-1])), c)
            j = order[start + np.arange(c.sum())]
            total += int(np.sum((np.abs(ay[j] - ay[i]) <= cs) & (np.abs(ax[j] - ax[i]) <= cs)))
    return total


def workload_bytes(kernel, S2, p, pairs_per_star, trace):
    """Algorithmic HBM bytes of all launches of `kernel` in one fit, from the
    unit costs b_f = b_i = 4 (stamp values, indices) and b_d = 8 (star
    vectors, normal-matrix blocks) over the live stars N of every trial, with
    E = N * pairs_per_star row entries (ordered pairs, each star with itself)
    among them and (E + N) / 2 distinct p x p blocks:
      k_gk     reads each distinct block once ((E + N) / 2 p^2 b_d), every row
               entry's word (neighbour, block slot or row length: E 2 b_i),
               the previous step's g and v and the x and w vectors (4 p N b_d)
               and writes g, v, x and w (4 p N b_d);
      k_hbuild reads every live star's stamp rows once (p N S^2 b_f) and the
               row entries (E b_i), and writes the distinct blocks.
    k_gk runs once per inner (Golub-Kahan) step, k_hbuild once per outer
    iteration (its first trial)."""
    total = 0.0
    seen = set()
    for r in trace:
        it, n_live, its = int(r[0]), r[5], r[6]
        E = n_live * pairs_per_star
        blocks = (E + n_live) / 2 * p * p * 8
        if kernel == "k_gk":
            total += its * (blocks + E * 8 + 8 * p * 8 * n_live)
        elif kernel == "k_hbuild" and it not in seen:
            seen.add(it)
            total += p * 4 * S2 * n_live + E * 4 + blocks
    return total


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary (FETCH_SIZE and WRITE_SIZE, scaled by a plain copy of known size),
    or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_crowded_epsf_4096x4096_200k_hbm_traffic.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return int(float(line.split(":", 1)[1]))
    except OSError:
        pass
    return None


def profile_tile(pkg, side, model, fits):
    """`fits` plain fits of one tile of the serving mix (inputs resident in
    HBM, row-major), for kernel traces of the small-tile path."""
    import torch
    abi = pkg._abi
    dev = torch.device("cuda:0")
    t = pkg.simulation.serving_tile(7, side, model)
    psf_model, mask = pkg.photometry._sim_psf_and_mask(t["psf"], t["fixed"])
    o = abi.sim_opts_default()
    o.model_rad = t["model_rad"]
    o.max_iter = t["max_iter"]
    o.free_mask = mask
    o.fixed_bkg = 0.0
    o.layout = abi.CP_LAYOUT_ROWMAJOR
    desc = psf_model.desc()
    src = t["sources"]
    dv = [torch.from_numpy(np.ascontiguousarray(a, np.float32).ravel()).to(dev) for a in (t["image"], t["inv_var"])]
    dv += [torch.from_numpy(np.ascontiguousarray(src[k])).to(dev) for k in ("y", "x", "flux")]
    n = t["n"]
    cols = {k: np.zeros(n) for k in ("y", "x", "flux", "flux_err")}
    r = abi.SimResult()
    for k, v in cols.items():
        setattr(r, k, v.ctypes.data_as(abi.f64p))
    ctx = pkg.Context(0)
    torch.cuda.synchronize()
    for i in range(fits):
        a = time.perf_counter()
        ctx.check(ctx.lib.cp_fit_simultaneous_device(
            ctx.handle, C.byref(desc), *[C.c_void_p(d.data_ptr()) for d in dv[:2]], side, side,
            *[C.c_void_p(d.data_ptr()) for d in dv[2:]], n, float(t["fit_rad"]), C.byref(o), C.byref(r)))
        log(f"[bench] profile tile {side}^2 {model}, {n} stars: fit {i} {(time.perf_counter() - a) * 1e3:.3f} ms")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-crop", type=int, default=2048)
    ap.add_argument("--side", type=int, default=4096)
    ap.add_argument("--stars", type=int, default=200_000)
    ap.add_argument("--profile-only", action="store_true", help="one fit, for rocprofv3 runs")
    ap.add_argument("--profile-fits", type=int, default=1,
                    help="fits run by --profile-only (the last one is the steady state)")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving headline (the c2 fit becomes it)")
    ap.add_argument("--serve-per-step", type=int, default=64, help="requests per timed serving step")
    ap.add_argument("--serve-open", type=int, default=1000, help="open-loop requests for p50 / p99 latency")
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--c2-steps", type=int, default=5, help="timed fits of the 4096^2 field")
    ap.add_argument("--lm-cutouts", type=int, default=1_000_000,
                    help="batched per-star LM cutouts (0 skips)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the built-in launch shapes instead of tuning them on the workload first")
    ap.add_argument("--no-plummer", action="store_true",
                    help="skip the 8192^2 field with 1e6 Plummer-clustered stars")
    ap.add_argument("--no-sequential", action="store_true", help="skip the sequential fit_all_stars section")
    ap.add_argument("--no-detection", action="store_true", help="skip the matched_filter section")
    ap.add_argument("--profile-detection", action="store_true",
                    help="one matched_filter call, for rocprofv3 runs")
    ap.add_argument("--profile-sequential", action="store_true",
                    help="one sequential 3-pass fit_all_stars, for rocprofv3 runs")
    ap.add_argument("--profile-lm", action="store_true",
                    help="one batched per-star LM call over the --lm-cutouts batch, for rocprofv3 runs")
    ap.add_argument("--profile-tile", type=int, default=0,
                    help="--profile-fits fits of one serving-mix tile of this side, for rocprofv3 runs")
    ap.add_argument("--profile-tile-model", default="gauss", choices=("gauss", "epsf"))
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus > 1 or world > 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is a design note only "
                                   "(the path would shard by image strips with a model_rad halo)",
                          "n_gpus": max(args.gpus, world)}))
        sys.exit(2)

    # the FP64 oracle on the whole 4096^2 field runs beside the GPU sections
    # (the CPU baseline, and the checker of the c2 fit)
    oracle_job = None
    if not args.no_cpu_baseline and args.side == 4096 and args.stars == 200_000 and not (
            args.profile_only or args.profile_detection or args.profile_sequential or args.profile_lm):
        oracle_job = start_oracle()

    import torch

    pkg = load_pkg()
    abi = pkg._abi
    sim = pkg.simulation
    if args.profile_tile:
        profile_tile(pkg, args.profile_tile, args.profile_tile_model, max(args.profile_fits, 1))
        return
    side, nstar = args.side, args.stars
    t0 = time.time()
    cfg = sim.config_epsf_crowded(seed=2, shape=(side, side), n=nstar, max_iter=10)
    log(f"[bench] generated {side}^2 field with {nstar} stars in {time.time() - t0:.1f}s")
    ny, nx = cfg["image"].shape
    dev = torch.device("cuda:0")
    d_img = torch.from_numpy(np.ascontiguousarray(cfg["image"].T).ravel()).to(dev)
    d_iv = torch.from_numpy(np.ascontiguousarray(cfg["inv_var"].T).ravel()).to(dev)
    src = cfg["sources"]
    d_y = torch.from_numpy(np.ascontiguousarray(src["y"])).to(dev)
    d_x = torch.from_numpy(np.ascontiguousarray(src["x"])).to(dev)
    d_f = torch.from_numpy(np.ascontiguousarray(src["flux"])).to(dev)
    torch.cuda.synchronize()

    ctx = pkg.Context(0)
    lib = ctx.lib
    psf = cfg["psf"]
    desc = psf.desc()
    n = len(src["y"])

    def make_opts(precision):
        o = abi.sim_opts_default()
        o.model_rad = cfg["model_rad"]
        o.max_iter = cfg["max_iter"]
        o.fixed_bkg = 0.0
        o.precision = precision
        return o

    def make_result(trace_cap=0):
        r = {k: np.zeros(n) for k in PER_STAR}
        r["converged"] = np.zeros(n, np.uint8)
        r["valid"] = np.zeros(n, np.uint8)
        r["n_iter"] = np.zeros(n, np.int32)
        r["trace"] = np.zeros(max(trace_cap, 1) * 8)
        s = abi.SimResult()
        for k in PER_STAR + ("trace",):
            setattr(s, k, r[k].ctypes.data_as(abi.f64p))
        s.converged = r["converged"].ctypes.data_as(abi.u8p)
        s.valid = r["valid"].ctypes.data_as(abi.u8p)
        s.n_iter = r["n_iter"].ctypes.data_as(abi.i32p)
        s.trace_cap = trace_cap
        return s, r

    def fit(o, s):
        rc = lib.cp_fit_simultaneous_device(ctx.handle, C.byref(desc), C.c_void_p(d_img.data_ptr()),
                                            C.c_void_p(d_iv.data_ptr()), ny, nx, C.c_void_p(d_y.data_ptr()),
                                            C.c_void_p(d_x.data_ptr()), C.c_void_p(d_f.data_ptr()), n,
                                            float(cfg["fit_rad"]), C.byref(o), C.byref(s))
        ctx.check(rc)

    o32 = make_opts(abi.CP_PREC_MIXED)
    s32, r32 = make_result(512)
    if args.profile_detection:
        detection_bench(pkg, ctx, cfg, d_img, d_iv, True, steps=0)
        log("[bench] profile-only matched_filter done")
        return
    if args.profile_sequential:
        sequential_bench(pkg, ctx, cfg, d_img, d_iv, True, steps=0)
        log("[bench] profile-only sequential fit done")
        return
    if args.profile_lm:
        lm_bench(pkg, ctx, args.lm_cutouts, True, False, steps=0)
        log("[bench] profile-only batched LM done")
        return
    if args.profile_only:
        # calibration: one plain device copy of 256 MiB (known read and write
        # bytes) for scaling the FETCH_SIZE / WRITE_SIZE counters
        a_cal = torch.ones(64 << 20, dtype=torch.float32, device=dev)
        b_cal = torch.empty_like(a_cal)
        b_cal.copy_(a_cal)
        torch.cuda.synchronize()
        del a_cal, b_cal
        for _ in range(max(args.profile_fits, 1)):
            fit(o32, s32)
        torch.cuda.synchronize()
        log(f"[bench] profile-only fits done: passes {s32.n_passes}")
        return
    tuning = None
    if not args.no_autotune:
        # launch shapes (blocks per CU of J v, J^T u and the renders) picked by
        # timing whole fits of this workload, before the timed region
        t0 = time.time()
        tuning, tbest = pkg.autotune.tune(ctx, lambda: fit(o32, s32), repeats=2)
        log(f"[bench] autotuned launch shapes in {time.time() - t0:.1f}s: {tuning} ({tbest * 1e3:.2f} ms)")
    for i in range(args.warmup):
        fit(o32, s32)
        log(f"[bench] warmup {i + 1}/{args.warmup} done")
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for i in range(args.c2_steps):
        fit(o32, s32)
    torch.cuda.synchronize()
    t_end = time.perf_counter()
    c2_ms = (t_end - t_start) * 1e3 / args.c2_steps
    c2_rate = n * args.c2_steps / (t_end - t_start)
    log(f"[bench] timed {args.c2_steps} fits: {c2_ms:.2f} ms/fit, {c2_rate:.4g} stars/s, passes {s32.n_passes}")
    trace32 = r32["trace"][: 8 * s32.trace_len].reshape(-1, 8).copy()

    # accuracy against the float64 device run on the same inputs
    o64 = make_opts(abi.CP_PREC_FP64)
    s64, r64 = make_result(512)
    t64 = time.perf_counter()
    fit(o64, s64)
    t64 = time.perf_counter() - t64
    g = (r32["valid"] > 0) & (r64["valid"] > 0)
    rel = np.abs(r32["flux"][g] - r64["flux"][g]) / np.abs(r64["flux"][g])
    dpos = np.maximum(np.abs(r32["y"][g] - r64["y"][g]), np.abs(r32["x"][g] - r64["x"][g]))
    within = float(np.mean((rel <= 1e-4) & (dpos <= 1e-3)))
    log(f"[bench] fp64 run {t64 * 1e3:.1f} ms; mixed vs fp64: {within * 100:.3f}% within 1e-4 flux / 1e-3 px, "
        f"median rel {np.median(rel):.2e}")
    truth_rel = np.abs(r32["flux"][g] / cfg["truth"][2][g] - 1)

    # instrumented fit: per-kernel HIP-event timing on the context stream
    lib.cp_profile(ctx.handle, 1)
    sp, rp = make_result(512)
    fit(o32, sp)
    buf = C.create_string_buffer(1 << 16)
    lib.cp_profile_report(ctx.handle, buf, len(buf))
    lib.cp_profile(ctx.handle, 0)
    kernels = {}
    for line in buf.value.decode().splitlines():
        name, cnt, ms = line.split()
        kernels[name] = (int(cnt), float(ms))
    log("[bench] kernels: " + ", ".join(f"{k} {c}x {t:.2f}ms" for k, (c, t) in sorted(kernels.items(), key=lambda kv: -kv[1][1])))
    tracep = rp["trace"][: 8 * sp.trace_len].reshape(-1, 8)
    R = int(math.ceil(cfg["fit_rad"]))
    S2 = (2 * R + 1) ** 2
    p = 3
    pairs = star_pairs(src["y"], src["x"], R)
    pps = pairs / n
    roof = None
    cand = {k: v for k, v in kernels.items() if k in ("k_gk", "k_hbuild")}
    if cand:
        # the dominant HBM-bound kernel (the renders and the fill are bound by
        # ePSF-evaluation issue, not HBM)
        dom = max(cand, key=lambda k: cand[k][1])
        cnt, tot_ms = cand[dom]
        bytes_total = workload_bytes(dom, S2, p, pps, tracep)
        avg_ms = tot_ms / cnt
        per_launch = bytes_total / cnt
        achieved = per_launch / (avg_ms * 1e-3) / 1e9
        roof = {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(dom),
                "avg_launch_us": round(avg_ms * 1e3, 2), "launches": cnt,
                "algorithmic_bytes_per_launch": int(per_launch),
                "workload": f"the {ny}x{nx} crowded ePSF tile ({n} stars, {pairs} overlapping star pairs), "
                            "instrumented fit"}
    c2 = {"workload": f"crowded ePSF field {ny}x{nx}, {n} stars, ePSF 101x101 os4, fit_rad 3, model_rad 12, "
                      f"10 outer GN iterations, LSQR 10 inner",
          "stars_per_s": round(c2_rate, 1), "ms_per_fit": round(c2_ms, 3), "fits_timed": args.c2_steps,
          "pixels_per_s": round(ny * nx / (c2_ms * 1e-3), 1), "outer_iterations": int(s32.n_passes),
          "star_pairs": pairs, "launch_shapes": tuning or "built-in defaults",
          "accuracy_vs_fp64_device": {"frac_within_1e-4_flux_1e-3_px": within,
                                      "median_rel_flux": float(np.median(rel)),
                                      "p99_rel_flux": float(np.quantile(rel, 0.99)),
                                      "median_rel_flux_vs_truth": float(np.median(truth_rel))},
          "fp64_ms_per_fit": round(t64 * 1e3, 2),
          "kernels_ms": {k: round(v[1], 3) for k, v in kernels.items()},
          "kernel_launches": {k: v[0] for k, v in kernels.items()}}
    r32_host = {k: r32[k].copy() for k in ("y", "x", "flux", "valid", "n_iter")}

    serving = None
    if not args.no_serving:
        serving = serving_bench(pkg, args.steps, args.warmup, args.serve_per_step, args.serve_open, args.serve_slots)
    batched_lm = (lm_bench(pkg, ctx, args.lm_cutouts, args.no_cpu_baseline, not args.no_autotune)
                  if args.lm_cutouts > 0 else None)
    detection = None
    if not args.no_detection:
        detection = detection_bench(pkg, ctx, cfg, d_img, d_iv, args.no_cpu_baseline)
    sequential = None
    if not args.no_sequential:
        sequential = sequential_bench(pkg, ctx, cfg, d_img, d_iv, args.no_cpu_baseline)

    plummer = None
    if not args.no_plummer:
        del d_img, d_iv, d_y, d_x, d_f
        torch.cuda.empty_cache()
        plummer = plummer_bench(pkg, ctx)

    cpu = None
    if oracle_job is not None:
        cpu, acc_oracle = finish_oracle(*oracle_job, r32_host)
        c2["accuracy_vs_fp64_oracle"] = acc_oracle
    elif not args.no_cpu_baseline:
        cpu = cpu_baseline(pkg, args.cpu_crop)

    if serving is not None:
        value = serving["served_stars_per_s"]
        head = {"metric": "stars fitted/sec + p50/p99 tile latency, served vs plain GPU at 1e-4 flux err",
                "value": round(value, 1), "unit": "stars/s", "n_gpus": 1, "steps": args.steps,
                "warmup": args.warmup, "ms_per_step": round(serving["ms_per_step"], 3)}
        cfg_out = {"workload": serving["workload"], "requests_per_step": serving["requests_per_step"],
                   "slots": serving["slots"], "requests_timed": serving["requests_timed"],
                   "stars_timed": serving["stars_timed"]}
        extra = {"latency": serving["latency"], "plain": serving["plain"],
                 "served_vs_plain": serving["served_vs_plain"],
                 "served_vs_plain_target": serving["served_vs_plain_target"],
                 "results_identical_to_plain": serving["results_identical_to_plain"],
                 "accuracy": serving["accuracy"]}
    else:
        head = {"metric": "stars fitted/sec (simultaneous PSF fit, crowded ePSF field)", "value": round(c2_rate, 1),
                "unit": "stars/s", "n_gpus": 1, "steps": args.c2_steps, "warmup": args.warmup,
                "ms_per_step": round(c2_ms, 3)}
        cfg_out = {"workload": c2["workload"]}
        extra = {}
    out = dict(head)
    out.update({"higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                "dtype": "f32 pixels / f64 reductions", "data": "synthetic (seeded simulate_image semantics)",
                "config": cfg_out})
    out.update(extra)
    out.update({"roofline": roof, "cpu_baseline": cpu, "c2_plain": c2, "batched_lm": batched_lm,
                "sequential": sequential, "detection": detection, "plummer_8192": plummer})
    print(json.dumps(out), flush=True)


def serving_bench(pkg, steps, warmup, per_step, n_open, slots, seed=5):
    """Serving workload (the headline): uneven image tiles (side 512 / 1024 /
    2048 / 4096, side^2 / 80 stars clamped to 1000..200000, Gaussian or ePSF
    model, drawn uniformly) with their images, weights and catalogues resident
    in HBM (C order, transposed on the device), 200 GB HBM budget.  A step is
    per_step requests submitted at once to the continuous-batching server and
    waited for; `steps` timed steps after `warmup`.  The plain arm fits the
    same requests one at a time on one context.  Latency: n_open requests with
    Poisson arrivals at 70 % of the served capacity.  Every served result is
    compared with the plain one, and the plain mixed-precision fit of every
    tile of the pool with its float64 fit."""
    import torch
    abi = pkg._abi
    sim = pkg.simulation
    rng = np.random.default_rng(seed)
    dev = torch.device("cuda:0")
    sides = (512, 1024, 2048, 4096)
    t0 = time.time()
    pool = {}
    for i, sd in enumerate(sides):
        for j, m in enumerate(("gauss", "epsf")):
            t = sim.serving_tile(seed * 100 + i * 2 + j, sd, m)
            model, mask = pkg.photometry._sim_psf_and_mask(t["psf"], t["fixed"])
            o = abi.sim_opts_default()
            o.model_rad = t["model_rad"]
            o.max_iter = t["max_iter"]
            o.free_mask = mask
            o.fixed_bkg = 0.0
            o.layout = abi.CP_LAYOUT_ROWMAJOR
            o64 = abi.sim_opts_default()
            C.memmove(C.byref(o64), C.byref(o), C.sizeof(o))
            o64.precision = abi.CP_PREC_FP64
            src = t["sources"]
            pool[(sd, m)] = dict(
                n=t["n"], ny=sd, nx=sd, fit_rad=float(t["fit_rad"]), desc=model.desc(), opts=o, opts64=o64,
                img=torch.from_numpy(np.ascontiguousarray(t["image"], np.float32).ravel()).to(dev),
                iv=torch.from_numpy(np.ascontiguousarray(t["inv_var"], np.float32).ravel()).to(dev),
                y=torch.from_numpy(np.ascontiguousarray(src["y"])).to(dev),
                x=torch.from_numpy(np.ascontiguousarray(src["x"])).to(dev),
                f=torch.from_numpy(np.ascontiguousarray(src["flux"])).to(dev))
    torch.cuda.synchronize()
    draw = lambda k: [(sides[rng.integers(4)], ("gauss", "epsf")[rng.integers(2)]) for _ in range(k)]
    step_keys = [draw(per_step) for _ in range(warmup + steps)]
    open_keys = draw(n_open)
    log(f"[serve] {len(pool)} tiles generated and uploaded in {time.time() - t0:.1f}s; {steps} x {per_step} "
        f"requests timed, {n_open} open-loop requests, {slots} slots")

    def new_out(n):
        arrs = {k: np.zeros(n) for k in ("y", "x", "flux", "bkg", "y_err", "x_err", "flux_err", "bkg_err", "chisq")}
        arrs.update(converged=np.zeros(n, np.uint8), valid=np.zeros(n, np.uint8), n_iter=np.zeros(n, np.int32))
        r = abi.SimResult()
        for k in ("y", "x", "flux", "bkg", "y_err", "x_err", "flux_err", "bkg_err", "chisq"):
            setattr(r, k, arrs[k].ctypes.data_as(abi.f64p))
        r.converged = arrs["converged"].ctypes.data_as(abi.u8p)
        r.valid = arrs["valid"].ctypes.data_as(abi.u8p)
        r.n_iter = arrs["n_iter"].ctypes.data_as(abi.i32p)
        return r, arrs

    ctx = pkg.Context(0)
    lib = ctx.lib

    def fit_plain(k, fp64=False):
        t = pool[k]
        r, arrs = new_out(t["n"])
        rc = lib.cp_fit_simultaneous_device(ctx.handle, C.byref(t["desc"]), C.c_void_p(t["img"].data_ptr()),
                                            C.c_void_p(t["iv"].data_ptr()), t["ny"], t["nx"],
                                            C.c_void_p(t["y"].data_ptr()), C.c_void_p(t["x"].data_ptr()),
                                            C.c_void_p(t["f"].data_ptr()), t["n"], t["fit_rad"],
                                            C.byref(t["opts64"] if fp64 else t["opts"]), C.byref(r))
        ctx.check(rc)
        return arrs

    # plain arm and the accuracy of each pool tile against its float64 fit
    ref = {k: fit_plain(k) for k in pool}
    acc_in, acc_n, worst = 0, 0, 0.0
    for k in pool:
        a, b = ref[k], fit_plain(k, fp64=True)
        g = (a["valid"] > 0) & (b["valid"] > 0)
        rel = np.abs(a["flux"][g] - b["flux"][g]) / np.abs(b["flux"][g])
        dp = np.maximum(np.abs(a["y"][g] - b["y"][g]), np.abs(a["x"][g] - b["x"][g]))
        w = sum(1 for q in step_keys[warmup:] for kk in q if kk == k)
        acc_in += w * np.sum((rel <= 1e-4) & (dp <= 1e-3))
        acc_n += w * g.sum()
        worst = max(worst, float(np.quantile(rel, 0.999)))
    plain_keys = [k for q in step_keys[warmup:] for k in q]
    lat_plain = []
    tp = time.perf_counter()
    for k in plain_keys:
        a = time.perf_counter()
        fit_plain(k)
        lat_plain.append(time.perf_counter() - a)
    tp = time.perf_counter() - tp
    stars = sum(pool[k]["n"] for k in plain_keys)

    def run_server(srv, keys, arrivals):
        keep, tickets = [], []
        t_first = time.perf_counter()
        for k, g in zip(keys, arrivals):
            dt = t_first + g - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            t = pool[k]
            r, arrs = new_out(t["n"])
            q = abi.Request()
            q.psf = C.pointer(t["desc"])
            q.image, q.inv_var = t["img"].data_ptr(), t["iv"].data_ptr()
            q.ny, q.nx = t["ny"], t["nx"]
            q.y0, q.x0, q.flux0 = t["y"].data_ptr(), t["x"].data_ptr(), t["f"].data_ptr()
            q.n_stars, q.fit_rad = t["n"], t["fit_rad"]
            q.opts = C.pointer(t["opts"])
            q.out = C.pointer(r)
            q.device_inputs = 1
            tk = C.c_uint64()
            rc = lib.cp_server_submit(srv, C.byref(q), C.byref(tk))
            if rc != 0:
                raise RuntimeError(lib.cp_server_last_error(srv).decode())
            keep.append((r, arrs, q))
            tickets.append(tk.value)
        stats = []
        for tk in tickets:
            st = abi.RequestStat()
            rc = lib.cp_server_wait(srv, tk, C.byref(st), None, 0)
            if rc != 0:
                raise RuntimeError(f"served fit failed ({rc})")
            stats.append(st)
        return [x[1] for x in keep], stats

    h = C.c_void_p()
    if lib.cp_server_create(C.byref(h), 0, 200 << 30, slots) != 0:
        raise RuntimeError("cp_server_create failed")
    same = True
    try:
        run_server(h, list(pool), np.zeros(len(pool)))   # warm the slots
        for q in step_keys[:warmup]:
            run_server(h, q, np.zeros(len(q)))
        step_s = []
        for q in step_keys[warmup:]:
            a = time.perf_counter()
            res, _ = run_server(h, q, np.zeros(len(q)))
            step_s.append(time.perf_counter() - a)
            same = same and all(all(np.array_equal(x[f], ref[k][f]) for f in ("y", "x", "flux", "flux_err", "n_iter"))
                                for x, k in zip(res, q))
        tc = float(np.sum(step_s))
        rate = 0.7 * len(plain_keys) / tc
        opened, ost = run_server(h, open_keys, np.cumsum(rng.exponential(1.0 / rate, len(open_keys))))
        same = same and all(np.array_equal(x["flux"], ref[k]["flux"]) for x, k in zip(opened, open_keys))
    finally:
        lib.cp_server_destroy(h)
    lat_open = np.array([st.t_end - st.t_submit for st in ost])
    t_open = max(st.t_end for st in ost) - min(st.t_submit for st in ost)
    open_stars = sum(pool[k]["n"] for k in open_keys)
    res = {
        "workload": "serving loop: image tiles resident in HBM, side uniform over {512, 1024, 2048, 4096}, "
                    "side^2/80 stars (1000..200000), CircularGaussian (fit_rad 3, model_rad 7) or the crowded-field "
                    "ePSF (fit_rad 3, model_rad 12) per request, max_iter 10, LSQR 10 inner, 200 GB HBM budget",
        "slots": slots, "requests_per_step": per_step, "requests_timed": len(plain_keys), "stars_timed": int(stars),
        "served_stars_per_s": stars / tc, "ms_per_step": tc * 1e3 / len(step_s),
        "plain": {"stars_per_s": round(stars / tp, 1), "p50_ms": round(float(np.median(lat_plain)) * 1e3, 2),
                  "p99_ms": round(float(np.quantile(lat_plain, 0.99)) * 1e3, 2)},
        "latency": {"requests": len(open_keys), "arrival_rate_per_s": round(rate, 2),
                    "stars_per_s": round(open_stars / t_open, 1),
                    "p50_ms": round(float(np.median(lat_open)) * 1e3, 2),
                    "p99_ms": round(float(np.quantile(lat_open, 0.99)) * 1e3, 2)},
        "served_vs_plain": round(tp / tc, 3), "served_vs_plain_target": 20.0,
        "results_identical_to_plain": bool(same),
        "accuracy": {"reference": "float64 device fit of each request's tile",
                     "frac_within_1e-4_flux_1e-3_px": float(acc_in / max(acc_n, 1)),
                     "p99.9_rel_flux_worst_tile": worst},
    }
    log(f"[serve] plain {stars / tp:.4g} stars/s; served {stars / tc:.4g} stars/s ({tp / tc:.2f}x); open loop "
        f"{len(open_keys)} requests at {rate:.1f} req/s: p50 {res['latency']['p50_ms']} ms, p99 "
        f"{res['latency']['p99_ms']} ms; identical {same}; within 1e-4 of fp64 "
        f"{res['accuracy']['frac_within_1e-4_flux_1e-3_px'] * 100:.3f}%")
    return res


def plummer_bench(pkg, ctx, steps=2):
    """Mixed-precision crowded fit at scale: 8192x8192 FP32 field with
    1,000,000 Plummer-clustered stars (scale radius 1500 px), the ePSF of the
    headline field, fit_rad 3, model_rad 12, 10 outer iterations; image,
    weights and catalogue resident in HBM, outputs copied to host inside the
    timed region; accuracy against the float64 device run on the same inputs."""
    import torch
    abi = pkg._abi
    lib = ctx.lib
    t0 = time.time()
    cfg = pkg.simulation.config_plummer()
    ny, nx = cfg["image"].shape
    log(f"[c4] generated {ny}^2 Plummer field in {time.time() - t0:.1f}s")
    dev = torch.device("cuda:0")
    d_img = torch.from_numpy(np.ascontiguousarray(cfg["image"].T).ravel()).to(dev)
    d_iv = torch.from_numpy(np.ascontiguousarray(cfg["inv_var"].T).ravel()).to(dev)
    src = cfg["sources"]
    d_y, d_x, d_f = (torch.from_numpy(np.ascontiguousarray(src[k])).to(dev) for k in ("y", "x", "flux"))
    desc = cfg["psf"].desc()
    n = len(src["y"])

    def run(prec):
        o = abi.sim_opts_default()
        o.model_rad = cfg["model_rad"]
        o.max_iter = cfg["max_iter"]
        o.fixed_bkg = 0.0
        o.precision = prec
        r = {k: np.zeros(n) for k in PER_STAR}
        r["valid"] = np.zeros(n, np.uint8)
        r["converged"] = np.zeros(n, np.uint8)
        r["n_iter"] = np.zeros(n, np.int32)
        sr = abi.SimResult()
        for k in PER_STAR:
            setattr(sr, k, r[k].ctypes.data_as(abi.f64p))
        sr.valid = r["valid"].ctypes.data_as(abi.u8p)
        sr.converged = r["converged"].ctypes.data_as(abi.u8p)
        sr.n_iter = r["n_iter"].ctypes.data_as(abi.i32p)
        torch.cuda.synchronize()
        t = time.perf_counter()
        rc = lib.cp_fit_simultaneous_device(ctx.handle, C.byref(desc), C.c_void_p(d_img.data_ptr()),
                                            C.c_void_p(d_iv.data_ptr()), ny, nx, C.c_void_p(d_y.data_ptr()),
                                            C.c_void_p(d_x.data_ptr()), C.c_void_p(d_f.data_ptr()), n,
                                            float(cfg["fit_rad"]), C.byref(o), C.byref(sr))
        ctx.check(rc)
        return r, sr, time.perf_counter() - t

    run(abi.CP_PREC_MIXED)
    dts = []
    for _ in range(steps):
        r32, s32, dt = run(abi.CP_PREC_MIXED)
        dts.append(dt)
    dt = float(np.mean(dts))
    r64, _, dt64 = run(abi.CP_PREC_FP64)
    g = (r32["valid"] > 0) & (r64["valid"] > 0)
    rel = np.abs(r32["flux"][g] - r64["flux"][g]) / np.abs(r64["flux"][g])
    dpos = np.maximum(np.abs(r32["y"][g] - r64["y"][g]), np.abs(r32["x"][g] - r64["x"][g]))
    within = float(np.mean((rel <= 1e-4) & (dpos <= 1e-3)))
    res = {"workload": f"{ny}x{nx} FP32 field, {n} stars Plummer-clustered (scale radius 1500 px), ePSF 101x101 "
                       f"os4, fit_rad 3, model_rad 12, 10 outer GN iterations",
           "stars_per_s": round(n / dt, 1), "ms_per_fit": round(dt * 1e3, 2),
           "pixels_per_s": round(ny * nx / dt, 1), "fp64_ms_per_fit": round(dt64 * 1e3, 2),
           "n_union": int(s32.n_union), "outer_iterations": int(s32.n_passes),
           "accuracy": {"reference": "float64 device run, same inputs", "frac_within_1e-4_flux_1e-3_px": within,
                        "median_rel_flux": float(np.median(rel)), "p99_rel_flux": float(np.quantile(rel, 0.99))}}
    log(f"[c4] {n / dt:.4g} stars/s, {dt * 1e3:.1f} ms per fit (fp64 {dt64 * 1e3:.1f} ms); "
        f"{within * 100:.2f}% within 1e-4 flux / 1e-3 px of fp64")
    return res


def lm_bench(pkg, ctx, n, no_cpu, autotune, steps=3):
    """Batched per-star Levenberg-Marquardt over uneven isolated-star cutouts
    (odd sides 9..31, CircularMoffat with FWHM U(1.8, 3.0), beta 2.5 fixed,
    free y, x, flux, bkg, Poisson noise), packed and resident in HBM.
    Algorithmic bytes: every accum! pass reads the cutout once (4 s^2 B), and
    a fit makes iterations + 1 passes."""
    import torch
    abi = pkg._abi
    lib = ctx.lib
    t0 = time.time()
    c = pkg.simulation.config_cutouts(seed=3, n=n)
    log(f"[lm] {n} cutouts generated in {time.time() - t0:.1f}s")
    dev = torch.device("cuda:0")
    d_pix = torch.from_numpy(c["pix"]).to(dev)
    d_off = torch.from_numpy(c["offsets"]).to(dev)
    d_ny = torch.from_numpy(c["ny"]).to(dev)
    d_nx = torch.from_numpy(c["nx"]).to(dev)
    d_th0 = torch.from_numpy(c["theta0"]).to(dev)
    d_th = torch.empty((n, 6), dtype=torch.float64, device=dev)
    d_cov = torch.empty((n, 16), dtype=torch.float64, device=dev)
    d_st = torch.empty(n * C.sizeof(abi.LMStat), dtype=torch.uint8, device=dev)
    model = pkg.CircularMoffatPSF(alpha=1.0, beta=c["beta"])
    desc = model.desc()
    opts = abi.LMOpts()
    lib.cp_lm_opts_default(C.byref(opts))
    torch.cuda.synchronize()

    def fit():
        rc = lib.cp_fit_star_batch_device(ctx.handle, C.byref(desc), C.c_void_p(d_pix.data_ptr()),
                                          C.c_void_p(d_off.data_ptr()), C.c_void_p(d_ny.data_ptr()),
                                          C.c_void_p(d_nx.data_ptr()), None, C.c_void_p(d_th0.data_ptr()),
                                          c["free_mask"], n, C.byref(opts), C.c_void_p(d_th.data_ptr()),
                                          C.c_void_p(d_cov.data_ptr()), C.c_void_p(d_st.data_ptr()))
        ctx.check(rc)

    fit()
    torch.cuda.synchronize()
    if steps == 0:
        return None
    shape = None
    if autotune:
        def run():
            fit()
            torch.cuda.synchronize()
        shape, _ = pkg.autotune.tune(ctx, run, space=pkg.autotune.LM_SPACE, repeats=2)
        log(f"[lm] autotuned lm_blocks_per_cu = {shape['lm_blocks_per_cu']}, lm_work_tile = {shape['lm_work_tile']}")
    t = time.perf_counter()
    for _ in range(steps):
        fit()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / steps
    st = np.frombuffer(d_st.cpu().numpy().tobytes(), dtype=np.dtype(
        [("minimum", "f8"), ("cost_init", "f8"), ("lambda_final", "f8"), ("chisq", "f8"), ("iterations", "i4"),
         ("converged", "u1"), ("x_converged", "u1"), ("f_converged", "u1"), ("g_converged", "u1"),
         ("status", "i4"), ("_pad", "i4")]))
    th = d_th.cpu().numpy()
    s2 = c["ny"].astype(np.float64) * c["nx"]
    passes = st["iterations"].astype(np.float64) + 1
    bytes_alg = float(np.sum(passes * 4 * s2))
    achieved = bytes_alg / dt / 1e9
    rel = np.abs(th[:, 4] / c["truth"][:, 4] - 1)
    res = {"workload": f"{n} isolated-star cutouts, odd sides 9..31 (mean {s2.mean():.1f} px), CircularMoffat "
                       f"FWHM U(1.8, 3.0) beta 2.5, free (y, x, flux, bkg), Poisson, unweighted",
           "fits_per_s": round(n / dt, 1), "ms_per_batch": round(dt * 1e3, 3),
           "lm_blocks_per_cu": shape["lm_blocks_per_cu"] if shape else "built-in default (8)",
           "lm_work_tile": shape["lm_work_tile"] if shape else "built-in default (1)",
           "pixels_per_s": round(float(s2.sum()) / dt, 1),
           "mean_iterations": round(float(st["iterations"].mean()), 3),
           "converged_frac": round(float(st["converged"].mean()), 5),
           "median_rel_flux_vs_truth": float(np.median(rel)),
           "roofline": {"bound": "latency (per-fit LM step: float64 reductions and solve; the cutout stays in "
                                 "L1/L2 across its accum! passes)",
                        "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4),
                        "bytes": "sum over fits of (iterations + 1) * 4 * s^2"}}
    log(f"[lm] {n / dt:.4g} fits/s, {dt * 1e3:.2f} ms/batch, mean iterations {st['iterations'].mean():.2f}, "
        f"{achieved:.0f} GB/s algorithmic")
    if not no_cpu:
        # FP64 oracle fit_star, one thread, over the leading cutouts of the
        # same batch until about 10 s have passed; the same cutouts give the
        # agreement of the GPU fits with the FP64 fitter
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from _support import oracle
        lib_o = oracle()
        o64 = abi.lm_opts_default()
        chunk, done, tc = 2000, 0, 0.0
        tho = np.zeros((n, 6))
        sto = (abi.LMStat * chunk)()
        ok_o = np.zeros(n, bool)
        while tc < 10.0 and done < n:
            k = min(chunk, n - done)
            off = c["offsets"][done:done + k] - c["offsets"][done]
            base = int(c["offsets"][done])
            t = time.perf_counter()
            lib_o.or_fit_star_batch(desc, c["pix"][base:].ctypes.data_as(abi.f32p),
                                    np.ascontiguousarray(off).ctypes.data_as(abi.i64p),
                                    c["ny"][done:].ctypes.data_as(abi.i32p), c["nx"][done:].ctypes.data_as(abi.i32p),
                                    None, c["theta0"][done:].ctypes.data_as(abi.f64p), c["free_mask"], k,
                                    C.byref(o64), tho[done:].ctypes.data_as(abi.f64p), None, sto, 1)
            tc += time.perf_counter() - t
            ok_o[done:done + k] = [sto[j].status == 0 for j in range(k)]
            done += k
        g = ok_o[:done] & (st["status"][:done] == 0)
        relc = np.abs(th[:done, 4][g] / tho[:done, 4][g] - 1)
        dposc = np.maximum(np.abs(th[:done, 0][g] - tho[:done, 0][g]), np.abs(th[:done, 1][g] - tho[:done, 1][g]))
        res["accuracy"] = {"reference": f"FP64 oracle fit_star on the first {done} cutouts",
                           "frac_within_1e-4_flux_1e-3_px": float(np.mean((relc <= 1e-4) & (dposc <= 1e-3))),
                           "median_rel_flux": float(np.median(relc))}
        res["cpu_baseline"] = {"value": round(done / tc, 1), "unit": "fits/s", "cores": 1, "kind": "port",
                               "sample": f"FP64 oracle fit_star, 1 thread, the first {done} cutouts of the batch, "
                                         f"{tc:.1f} s"}
        log(f"[lm] cpu baseline {done / tc:.1f} fits/s ({done} cutouts, {tc:.1f} s); GPU vs FP64 oracle "
            f"{res['accuracy']['frac_within_1e-4_flux_1e-3_px'] * 100:.3f}% within 1e-4 flux / 1e-3 px")
        # the same fitter with OpenMP over cutouts on every host thread this
        # process may use, over a sample about as long again
        nth = int(os.environ.get("OMP_NUM_THREADS") or len(os.sched_getaffinity(0)))
        k = min(n, max(chunk, int(done / tc * nth * 5.0)))
        sto_all = (abi.LMStat * k)()
        tho_all = np.zeros((k, 6))
        t = time.perf_counter()
        lib_o.or_fit_star_batch(desc, c["pix"].ctypes.data_as(abi.f32p), c["offsets"].ctypes.data_as(abi.i64p),
                                c["ny"].ctypes.data_as(abi.i32p), c["nx"].ctypes.data_as(abi.i32p), None,
                                c["theta0"].ctypes.data_as(abi.f64p), c["free_mask"], k, C.byref(o64),
                                tho_all.ctypes.data_as(abi.f64p), None, sto_all, nth)
        ta = time.perf_counter() - t
        res["cpu_baseline_all_cores"] = {
            "value": round(k / ta, 1), "unit": "fits/s", "cores": nth, "kind": "port",
            "sample": f"FP64 oracle fit_star, OpenMP over cutouts on {nth} threads, the first {k} cutouts of the "
                      f"batch, {ta:.1f} s",
            "same_as_one_thread": bool(np.array_equal(tho_all[:min(k, done)], tho[:min(k, done)]))}
        log(f"[lm] cpu all cores ({nth} threads): {k / ta:.1f} fits/s ({k} cutouts, {ta:.1f} s)")
    return res


def detection_bench(pkg, ctx, cfg, d_img, d_iv, no_cpu, steps=5):
    """matched_filter (zero-sum kernel, inv_var weights, sigma 5) on the
    crowded field with the fitting ePSF rendered as the kernel, image and
    weights resident in HBM: pixels per second.  The fused pass of the three
    2-D correlations (image, weighted image, weights with the squared kernel)
    dominates: 6 kr kc float64 flops per pixel."""
    abi = pkg._abi
    lib = ctx.lib
    ny, nx = cfg["image"].shape
    kern = pkg.render(cfg["psf"], ctx=ctx)
    kr, kc = kern.shape
    kv = np.asfortranarray(kern).ravel(order="F")
    cap = 1 << 20
    py, px = np.zeros(cap, np.int32), np.zeros(cap, np.int32)
    ps, pf = np.zeros(cap), np.zeros(cap)
    r = abi.MFResult()
    r.peaks_cap = cap
    r.peak_y, r.peak_x = py.ctypes.data_as(abi.i32p), px.ctypes.data_as(abi.i32p)
    r.peak_significance, r.peak_flux = ps.ctypes.data_as(abi.f64p), pf.ctypes.data_as(abi.f64p)

    def run():
        ctx.check(lib.cp_matched_filter_device(ctx.handle, C.c_void_p(d_img.data_ptr()), C.c_void_p(d_iv.data_ptr()),
                                               ny, nx, kv.ctypes.data_as(abi.f64p), kr, kc, 1, 5.0,
                                               abi.CP_BORDER_REPLICATE, C.byref(r)))

    run()
    if steps == 0:
        return None
    t = time.perf_counter()
    for _ in range(steps):
        run()
    dt = (time.perf_counter() - t) / steps
    lib.cp_profile(ctx.handle, 1)
    run()
    buf = C.create_string_buffer(1 << 16)
    lib.cp_profile_report(ctx.handle, buf, len(buf))
    lib.cp_profile(ctx.handle, 0)
    kernels = {}
    for line in buf.value.decode().splitlines():
        name, cnt, ms = line.split()
        kernels[name] = (int(cnt), float(ms))
    kcor = kernels.get("k_mf2d", (0, float("nan")))
    flops = 6.0 * kr * kc * ny * nx
    achieved = flops / (kcor[1] / kcor[0] * 1e-3) / 1e12 if kcor[0] else None
    n = int(r.n_peaks)
    truth_n = len(cfg["truth"][0])
    res = {"workload": f"matched_filter on the crowded ePSF field {ny}x{nx}: kernel = the ePSF rendered "
                       f"({kr}x{kc}), zero-sum, inv_var weights, sigma 5, replicate border",
           "pixels_per_s": round(ny * nx / dt, 1), "ms_per_call": round(dt * 1e3, 3), "peaks": n,
           "stars_in_field": truth_n, "kernels_ms": {k: round(v[1], 3) for k, v in kernels.items()},
           "roofline": {"bound": "fp64 (fused 2-D correlations, LDS-staged)", "kernel": "k_mf2d",
                        "achieved": round(achieved, 2) if achieved else None, "peak": 78.6, "unit": "TFLOP/s",
                        "frac": round(achieved / 78.6, 4) if achieved else None,
                        "flops": "6 kr kc per pixel (three correlations)"}}
    log(f"[det] {ny * nx / dt:.4g} px/s, {dt * 1e3:.2f} ms per call, {n} peaks; k_mf2d {kcor[0]}x "
        f"{kcor[1]:.2f} ms ({achieved or 0:.2f} TFLOP/s)")
    if not no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_detect
        crop = 2048
        img = cfg["image"][:crop, :crop].astype(np.float64)
        iv = cfg["inv_var"][:crop, :crop].astype(np.float64)
        t = time.perf_counter()
        oracle_detect.matched_filter(img, kern, inv_var=iv, sigma=5.0)
        tc = time.perf_counter() - t
        res["cpu_baseline"] = {"value": round(crop * crop / tc, 1), "unit": "pixels/s", "cores": 1, "kind": "port",
                               "sample": f"FP64 oracle matched_filter, 1 thread, {crop}x{crop} crop of the same "
                                         f"field and kernel, {tc:.1f} s"}
        log(f"[det] cpu baseline {crop * crop / tc:.4g} px/s ({tc:.1f} s)")
    return res


def sequential_bench(pkg, ctx, cfg, d_img, d_iv, no_cpu, steps=3):
    """Sequential multi-pass fit_all_stars (3 passes, fit_star defaults) over
    the same crowded ePSF field, inputs resident in HBM: stars fitted per
    second, with the FP64 oracle on a crop as the CPU figure."""
    import torch
    abi = pkg._abi
    lib = ctx.lib
    psf = cfg["psf"]
    src = cfg["sources"]
    n = len(src["y"])
    ny, nx = cfg["image"].shape
    th0 = np.zeros((n, 4))
    th0[:, 0], th0[:, 1], th0[:, 2] = src["y"], src["x"], src["flux"]
    d_th = torch.from_numpy(th0).to(torch.device("cuda:0"))
    names = psf.props
    fixv = np.array([0.0 if k == "bkg" else np.nan for k in names])
    mask = sum(1 << k for k, nme in enumerate(names) if nme != "bkg")
    o = abi.LMOpts()
    lib.cp_lm_opts_default(C.byref(o))
    desc = psf.desc()
    cols = {k: np.zeros(n) for k in ("y", "x", "flux", "bkg", "flux_err", "chisq", "qfit", "crowding")}
    conv = np.zeros(n, np.uint8)
    valid = np.zeros(n, np.uint8)
    nit = np.zeros(n, np.int32)
    r = abi.SimResult()
    for k, v in cols.items():
        setattr(r, k, v.ctypes.data_as(abi.f64p))
    r.converged = conv.ctypes.data_as(abi.u8p)
    r.valid = valid.ctypes.data_as(abi.u8p)
    r.n_iter = nit.ctypes.data_as(abi.i32p)
    torch.cuda.synchronize()

    def fit(prec):
        rc = lib.cp_fit_all_stars_device(ctx.handle, C.byref(desc), C.c_void_p(d_img.data_ptr()),
                                         C.c_void_p(d_iv.data_ptr()), ny, nx, C.c_void_p(d_th.data_ptr()), n,
                                         float(cfg["fit_rad"]), 3, mask, fixv.ctypes.data_as(abi.f64p), C.byref(o),
                                         0.0, prec, C.byref(r))
        ctx.check(rc)

    if steps == 0:
        fit(abi.CP_PREC_MIXED)
        return None
    fit(abi.CP_PREC_FP64)
    f64 = {k: v.copy() for k, v in cols.items()}
    v64 = valid.astype(bool)
    fit(abi.CP_PREC_MIXED)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(steps):
        fit(abi.CP_PREC_MIXED)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / steps
    lib.cp_profile(ctx.handle, 1)
    fit(abi.CP_PREC_MIXED)
    buf = C.create_string_buffer(1 << 16)
    lib.cp_profile_report(ctx.handle, buf, len(buf))
    lib.cp_profile(ctx.handle, 0)
    kernels = {}
    for line in buf.value.decode().splitlines():
        name, cnt, ms = line.split()
        kernels[name] = (int(cnt), float(ms))
    g = v64 & valid.astype(bool)
    rel = np.abs(cols["flux"][g] - f64["flux"][g]) / np.abs(f64["flux"][g])
    dpos = np.maximum(np.abs(cols["y"][g] - f64["y"][g]), np.abs(cols["x"][g] - f64["x"][g]))
    sig = np.abs(cols["flux"][g] - f64["flux"][g]) <= 0.02 * f64["flux_err"][g]
    within = float(np.mean(((rel <= 1e-4) | sig) & (dpos <= 1e-3)))
    # HBM bytes the fit kernel must move: each star's box of the float64
    # residual read and written and the float32 inverse variance read, per pass
    W = 2 * int(math.ceil(cfg["fit_rad"])) + 2   # box side for a non-integer centre
    box_bytes = 3.0 * n * W * W * (8 + 8 + 4)
    km = kernels.get("k_mp_level", (0, float("nan")))
    achieved = box_bytes / (km[1] * 1e-3) / 1e9 if km[0] else None
    res = {"workload": f"sequential fit_all_stars, 3 passes, on the crowded ePSF field {ny}x{nx} with {n} stars "
                       f"(fit_rad 3, bkg fixed at 0, fit_star defaults: max_iter 200)",
           "stars_per_s": round(n / dt, 1), "ms_per_fit": round(dt * 1e3, 2),
           "valid": int(valid.sum()), "mean_iterations_per_star": round(float(nit.mean()), 3),
           "iterations_per_star_p50_p99_max": [int(np.percentile(nit, 50)), int(np.percentile(nit, 99)),
                                               int(nit.max())],
           "kernels_ms": {k: round(v[1], 3) for k, v in kernels.items()},
           "fit_kernel_launches": km[0],
           "accuracy": {"reference": "float64 device run, same inputs",
                        "frac_within_1e-4_flux_or_0.02_sigma_and_1e-3_px": within,
                        "median_rel_flux": float(np.median(rel)) if rel.size else None},
           "roofline": {"bound": "latency (per-star LM iterations in LDS)", "achieved": round(achieved, 1)
                        if achieved else None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 5) if achieved else None,
                        "bytes": "3 passes x n x (2 ceil(fit_rad) + 2)^2 x 20 B"}}
    log(f"[seq] {n / dt:.4g} stars/s, {dt * 1e3:.1f} ms per 3-pass fit, k_mp_level {km[0]}x {km[1]:.1f} ms, "
        f"mixed vs fp64 {within * 100:.3f}%")
    if not no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_fit
        crop = 4096
        nst = int(round(n * crop * crop / (ny * nx)))
        c2 = pkg.simulation.config_epsf_crowded(seed=7, shape=(crop, crop), n=nst, max_iter=10)
        s2 = c2["sources"]
        t = time.perf_counter()
        oracle_fit.fit_all(psf, c2["image"].astype(np.float64), dict(y=s2["y"], x=s2["x"], flux=s2["flux"]), 3,
                           n_passes=3, fixed=dict(bkg=0.0), inv_var=c2["inv_var"].astype(np.float64))
        tc = time.perf_counter() - t
        res["cpu_baseline"] = {"value": round(nst / tc, 2), "unit": "stars/s", "cores": 1, "kind": "port",
                               "sample": f"FP64 oracle fit_all_stars, 1 thread, a {crop}x{crop} field of the "
                                         f"same workload (other seed, {nst} stars, 3 passes), {tc:.1f} s"}
        log(f"[seq] cpu baseline {nst / tc:.1f} stars/s ({tc:.1f} s)")
    return res


def oracle_worker(out_path):
    """Child process of the bench: the FP64 oracle (single thread) fits the
    whole crowded 4096^2 ePSF field (the largest tile of the serving mix and
    the c2 workload, regenerated from its seed), timed, and saves its
    per-star results for the accuracy comparison."""
    pkg = load_pkg()
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_fit
    cfg = pkg.simulation.config_epsf_crowded(seed=2, shape=(4096, 4096), n=200_000, max_iter=10)
    src = cfg["sources"]
    t = time.perf_counter()
    r = oracle_fit.fit(cfg["psf"], cfg["image"].astype(np.float64), src["y"], src["x"], src["flux"], cfg["fit_rad"],
                       inv_var=cfg["inv_var"], model_rad=cfg["model_rad"], max_iter=cfg["max_iter"])
    dt = time.perf_counter() - t
    np.savez(out_path, seconds=dt, n_passes=r.n_passes, **{k: getattr(r, k) for k in
             ("y", "x", "flux", "y_err", "x_err", "flux_err", "valid", "n_iter", "converged")})


def start_oracle():
    import tempfile
    path = os.path.join(tempfile.gettempdir(), f"crowdphot_oracle_{os.getpid()}.npz")
    proc = subprocess.Popen([sys.executable, os.path.abspath(__file__), "--oracle-worker", path],
                            stdout=subprocess.DEVNULL, stderr=subprocess.PIPE)
    return proc, path


def finish_oracle(proc, path, r32, timeout=900):
    """Waits for the oracle child; returns (cpu_baseline record, accuracy of
    the mixed device fit r32 against it, with the stars outside the bar)."""
    try:
        _, err = proc.communicate(timeout=timeout)
    except subprocess.TimeoutExpired:
        proc.kill()
        log("[bench] oracle run timed out")
        return None, None
    if proc.returncode != 0 or not os.path.exists(path):
        log(f"[bench] oracle run failed: {err.decode()[-400:]}")
        return None, None
    o = dict(np.load(path))
    os.remove(path)
    dt, n = float(o["seconds"]), len(o["flux"])
    model = ""
    try:
        model = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        model = [l.split(":", 1)[1].strip() for l in model.splitlines() if l.startswith("Model name")][0]
    except Exception:
        pass
    cpu = {"value": round(n / dt, 2), "unit": "stars/s", "cores": 1, "kind": "port",
           "sample": f"FP64 oracle, 1 thread, the whole 4096x4096 crowded ePSF field ({n} stars, 10 outer "
                     f"iterations; the largest tile of the serving mix), {dt:.1f} s, run beside the GPU sections",
           "cpu_model": model, "nproc": os.cpu_count(),
           "cpus_usable": len(os.sched_getaffinity(0)), "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}
    g = (r32["valid"] > 0) & (o["valid"] > 0)
    idx = np.nonzero(g)[0]
    rel = np.abs(r32["flux"][g] - o["flux"][g]) / np.abs(o["flux"][g])
    dp = np.maximum(np.abs(r32["y"][g] - o["y"][g]), np.abs(r32["x"][g] - o["x"][g]))
    ok = (rel <= 1e-4) & (dp <= 1e-3)
    out = idx[~ok]
    sf = np.abs(r32["flux"][out] - o["flux"][out]) / o["flux_err"][out]
    sp = np.maximum(np.abs(r32["y"][out] - o["y"][out]) / o["y_err"][out],
                    np.abs(r32["x"][out] - o["x"][out]) / o["x_err"][out])
    same_iter = r32["n_iter"][out] == o["n_iter"][out]
    worst = out[np.argsort(-sf)[:10]]
    acc = {"reference": "FP64 oracle (CPU), same inputs", "stars": int(g.sum()),
           "validity_equal": bool(np.array_equal(r32["valid"] > 0, o["valid"] > 0)),
           "frac_within_1e-4_flux_1e-3_px": float(ok.mean()), "median_rel_flux": float(np.median(rel)),
           "p99_rel_flux": float(np.quantile(rel, 0.99)), "out_of_bar": int(out.size),
           "out_of_bar_dflux_over_sigma": {"median": float(np.median(sf)) if out.size else 0.0,
                                           "p99": float(np.quantile(sf, 0.99)) if out.size else 0.0,
                                           "max": float(sf.max()) if out.size else 0.0},
           "out_of_bar_dpos_over_sigma_max": float(sp.max()) if out.size else 0.0,
           "out_of_bar_live_at_max_iter_both_sides": int(np.sum((r32["n_iter"][out] == 10) & (o["n_iter"][out] == 10))),
           "out_of_bar_same_freeze_iteration": int(same_iter.sum()),
           "worst": [{"star": int(i), "flux": float(o["flux"][i]), "dflux_over_sigma":
                      float(abs(r32["flux"][i] - o["flux"][i]) / o["flux_err"][i]),
                      "freeze_iteration_gpu": int(r32["n_iter"][i]), "freeze_iteration_oracle": int(o["n_iter"][i])}
                     for i in worst]}
    log(f"[bench] oracle: {n} stars in {dt:.1f}s ({n / dt:.1f} stars/s); mixed device fit vs oracle: "
        f"{ok.mean() * 100:.3f}% within 1e-4 flux / 1e-3 px, {out.size} outside (|dflux|/sigma max "
        f"{acc['out_of_bar_dflux_over_sigma']['max']:.3g})")
    return cpu, acc


def cpu_baseline(pkg, crop):
    """FP64 oracle (single thread) on a crop of the same workload at the same
    star density: stars fitted per second."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    try:
        import oracle_fit
    except Exception as e:  # pragma: no cover
        log(f"[bench] cpu baseline unavailable: {e}")
        return None
    sim = pkg.simulation
    nst = int(round(200_000 * crop * crop / 4096 ** 2))
    cfg = sim.config_epsf_crowded(seed=7, shape=(crop, crop), n=nst, max_iter=10)
    src = cfg["sources"]
    t = time.perf_counter()
    r = oracle_fit.fit(cfg["psf"], cfg["image"].astype(np.float64), src["y"], src["x"], src["flux"], cfg["fit_rad"],
                       inv_var=cfg["inv_var"], model_rad=cfg["model_rad"], max_iter=cfg["max_iter"])
    dt = time.perf_counter() - t
    cores = 1
    model = ""
    try:
        model = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        model = [l.split(":", 1)[1].strip() for l in model.splitlines() if l.startswith("Model name")][0]
    except Exception:
        pass
    log(f"[bench] cpu baseline: {nst} stars in {dt:.1f}s ({nst / dt:.1f} stars/s), passes {r.n_passes}")
    return {"value": round(nst / dt, 2), "unit": "stars/s", "cores": cores, "kind": "port",
            "sample": f"FP64 oracle, 1 thread, {crop}x{crop} crop of the same workload ({nst} stars, same "
                      f"density, 10 outer iterations), {dt:.1f} s", "cpu_model": model}


if __name__ == "__main__":
    if len(sys.argv) == 3 and sys.argv[1] == "--oracle-worker":
        oracle_worker(sys.argv[2])
    else:
        main()
