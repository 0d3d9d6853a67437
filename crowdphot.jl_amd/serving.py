"""Continuous-batching server for the simultaneous fit (cp_server_* in
include/crowdphot_hip.h).

Requests (image tiles with their catalogues) are submitted at any time; the
library admits them in arrival order as soon as a worker slot is free and
their HBM estimate fits the budget, and the admitted fits overlap on the GPU.
Each result equals what the same request gets alone.  There is no counterpart
in the reference, which fits one image per call
(fit_all_stars_simultaneous, src/photometry/psf_photometry_simultaneous.jl:953).
"""
import ctypes as C

from . import _abi
from .photometry import ArgumentError, _SimCall


class Server:
    """``Server(device=0, hbm_budget=0, slots=4)``; ``submit(...)`` returns a
    ticket, ``wait(ticket)`` returns ``(MultiPassPhotResult, stat)`` where stat
    holds the submit / start / end times (seconds) and the slot."""

    def __init__(self, device=0, hbm_budget=0, slots=4):
        self.lib = _abi.load_library()
        h = C.c_void_p()
        rc = self.lib.cp_server_create(C.byref(h), int(device), int(hbm_budget), int(slots))
        if rc != _abi.CP_OK:
            raise RuntimeError(f"cp_server_create failed ({rc}): no usable HIP device {device}")
        self.handle = h
        self._pending = {}

    def submit(self, image, psf, sources, fit_rad, **kw):
        """Queue one fit; keyword arguments as fit_all_stars_simultaneous."""
        call = _SimCall(image, psf, sources, fit_rad, **kw)
        req = _abi.Request()
        req.psf = C.pointer(call.desc)
        req.image = call.im.ctypes.data
        req.inv_var = None if call.iv is None else call.iv.ctypes.data
        req.ny, req.nx = call.ny, call.nx
        req.y0, req.x0, req.flux0 = call.y0.ctypes.data, call.x0.ctypes.data, call.f0.ctypes.data
        req.n_stars = call.n
        req.fit_rad = call.fit_rad
        req.opts = C.pointer(call.opts)
        req.out = C.pointer(call.result)
        req.device_inputs = 0
        t = C.c_uint64()
        rc = self.lib.cp_server_submit(self.handle, C.byref(req), C.byref(t))
        if rc != _abi.CP_OK:
            msg = self.lib.cp_server_last_error(self.handle).decode()
            if rc == _abi.CP_EARG:
                raise ArgumentError(msg)
            if rc == _abi.CP_ENOMEM:
                raise MemoryError(msg)
            raise RuntimeError(msg)
        self._pending[t.value] = (call, req)   # keeps every array alive until wait
        return t.value

    def done(self, ticket):
        d = C.c_int32()
        if self.lib.cp_server_poll(self.handle, int(ticket), C.byref(d)) != _abi.CP_OK:
            raise ArgumentError(f"unknown ticket {ticket}")
        return bool(d.value)

    def wait(self, ticket):
        st = _abi.RequestStat()
        msg = C.create_string_buffer(512)
        rc = self.lib.cp_server_wait(self.handle, int(ticket), C.byref(st), msg, 512)
        call, _ = self._pending.pop(ticket)
        if rc != _abi.CP_OK:
            text = msg.value.decode()
            if rc == _abi.CP_EARG:
                raise ArgumentError(text)
            if rc == _abi.CP_ENOMEM:
                raise MemoryError(text)
            raise RuntimeError(f"libcrowdphot_hip error {rc}: {text}")
        stat = dict(slot=st.slot, t_submit=st.t_submit, t_start=st.t_start, t_end=st.t_end,
                    latency=st.t_end - st.t_submit, hbm_bytes=st.hbm_bytes, hbm_charged=st.hbm_charged,
                    hbm_peak=st.hbm_peak, hbm_resident_peak=st.hbm_resident_peak)
        return call.finish(), stat

    def close(self):
        if getattr(self, "handle", None):
            for t in list(self._pending):
                self.wait(t)
            self.lib.cp_server_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
