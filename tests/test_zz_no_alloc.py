"""cp_fit_star_batch_device_noalloc against cp_fit_star_batch_device on the
same device inputs, through the C ABI: identical results on two calls in a
row (the second smaller), for an analytic model and a four-node ePSF, no
device block taken from the context, and PSFs too large for the scratch
array rejected."""
import ctypes as C

import numpy as np
import pytest

from _support import DeviceArray, load_pkg

pkg = load_pkg()
abi = pkg._abi
sim = pkg.simulation

pytestmark = pytest.mark.gpu


def _epsf4():
    e = pkg.moffat_epsf(fwhm=2.2, beta=2.5, size=41, oversampling=4)
    e2 = pkg.moffat_epsf(fwhm=2.7, beta=3.0, size=41, oversampling=4)
    return pkg.GriddedPSFModel([e, e2, e2, e], [1.0, 1.0, 192.0, 192.0], [1.0, 160.0, 1.0, 160.0])


def _epsf_cutouts(psf, n, seed):
    rng = np.random.default_rng(seed)
    cuts, th0 = [], []
    for i in range(n):
        s = int(rng.choice([11, 13, 15, 17]))
        y, x = (s + 1) / 2 + rng.uniform(-0.5, 0.5), (s + 1) / 2 + rng.uniform(-0.5, 0.5)
        f, b = rng.uniform(500, 5000), rng.uniform(5, 50)
        c = sim.render_sources(psf, (s, s), [y], [x], [f], 40) + b
        cuts.append(sim.add_noise(c, seed + i).astype(np.float32))
        th0.append([y + rng.normal(0, 0.2), x + rng.normal(0, 0.2), f * rng.uniform(0.8, 1.2), b])
    ny = np.array([c.shape[0] for c in cuts], np.int32)
    off = np.zeros(n, np.int64)
    off[1:] = np.cumsum(ny.astype(np.int64) ** 2)[:-1]
    pix = np.concatenate([c.T.ravel() for c in cuts]).astype(np.float32)
    return dict(pix=pix, offsets=off, ny=ny, nx=ny.copy(), theta0=np.array(th0), free_mask=0b1111)


def _moffat_cutouts(n, seed):
    c = sim.config_cutouts(seed=seed, n=n)
    return dict(pix=c["pix"], offsets=c["offsets"], ny=c["ny"], nx=c["nx"], theta0=c["theta0"],
                free_mask=c["free_mask"], beta=c["beta"])


def _run(ctx, fn, desc, c):
    n = len(c["ny"])
    npf = c["theta0"].shape[1]
    p = bin(c["free_mask"]).count("1")
    d = {k: DeviceArray(np.ascontiguousarray(c[k])) for k in ("pix", "offsets", "ny", "nx", "theta0")}
    th = DeviceArray(nbytes=n * npf * 8)
    cov = DeviceArray(nbytes=n * p * p * 8)
    st = DeviceArray(nbytes=n * C.sizeof(abi.LMStat))
    o = abi.lm_opts_default()
    rc = fn(ctx.handle, C.byref(desc), d["pix"].ptr, d["offsets"].ptr, d["ny"].ptr, d["nx"].ptr, None,
            d["theta0"].ptr, c["free_mask"], n, C.byref(o), th.ptr, cov.ptr, st.ptr)
    assert rc == abi.CP_OK, ctx.lib.cp_last_error(ctx.handle).decode()
    return th.get(np.float64, (n, npf)), cov.get(np.float64, (n, p, p)), st.get(np.uint8)


def _subset(c, m):
    out = dict(c)
    out["ny"], out["nx"], out["offsets"], out["theta0"] = c["ny"][:m], c["nx"][:m], c["offsets"][:m], c["theta0"][:m]
    end = int(c["offsets"][m - 1] + c["ny"][m - 1] * c["nx"][m - 1])
    out["pix"] = c["pix"][:end]
    return out


@pytest.mark.parametrize("kind", ["moffat", "epsf4"])
def test_noalloc_equals_allocating_entry_point(kind):
    print(f"case: no-alloc batched LM, {kind}", flush=True)
    ctx = pkg.Context(0)
    lib = ctx.lib
    if kind == "moffat":
        c = _moffat_cutouts(3000, 11)
        model = pkg.CircularMoffatPSF(alpha=1.0, beta=c["beta"])
    else:
        model = _epsf4()
        c = _epsf_cutouts(model, 300, 21)
    desc = model.desc()
    for m in (len(c["ny"]), len(c["ny"]) // 3):   # a second call with a smaller input
        cm = _subset(c, m)
        ref = _run(ctx, lib.cp_fit_star_batch_device, desc, cm)
        ctx.set_budget(0)   # restarts the context's peak block count
        got = _run(ctx, lib.cp_fit_star_batch_device_noalloc, desc, cm)
        assert ctx.mem_stats()["peak"] == 0, "the allocation-free call took a device block"
        for a, b in zip(ref, got):
            np.testing.assert_array_equal(a, b)
        ok = np.frombuffer(got[2].tobytes(), dtype=abi.LMStat)["status"] == 0
        assert ok.mean() > 0.99


def test_noalloc_rejects_psf_beyond_scratch():
    ctx = pkg.Context(0)
    side = 1451   # 1451^2 float32 node grid: 8.4 MB > 8 MiB of scratch
    big = pkg.GriddedPSFModel([pkg.ImagePSF(np.ones((side, side)), oversampling=4)], [0.0], [0.0])
    c = _moffat_cutouts(4, 3)
    c["theta0"] = np.ascontiguousarray(c["theta0"][:, [0, 1, 4, 5]])
    c["free_mask"] = 0b1111
    with pytest.raises(AssertionError, match="scratch"):
        _run(ctx, ctx.lib.cp_fit_star_batch_device_noalloc, big.desc(), c)
