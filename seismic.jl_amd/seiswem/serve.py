"""Serving loop (continuous batching of migration requests) over the C ABI (swem_server_*,
swem_batch_*).  The reference has no serving path; its driver starts one shotwem process per
shot (src/Imaging/ShotProfileWEM.jl:104-128, 288).  A request is the in-memory ShotProfileWEM
of a few shots over one velocity (nangx = nangy = 1)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check, load
from .wem import WemGeometry


class Server:
    def __init__(self, device_bytes: int = 0, pinned_bytes: int = 0, streams: int = 4, max_inflight: int = 8,
                 max_pending: int = 4096, skip_window_ms: float = 500.0, device: int = 0, max_fuse: int = 32):
        """max_fuse: single-shot requests over the same velocity, wavelet and geometry that are
        queued together run as the shots of one launch (at most max_fuse; 1 = never)."""
        c = _lib.ServerConfig()
        load().swem_server_config_init(ctypes.byref(c))
        c.device, c.device_bytes, c.pinned_bytes = device, device_bytes, pinned_bytes
        c.streams, c.max_inflight, c.max_pending, c.skip_window_ms = streams, max_inflight, max_pending, skip_window_ms
        c.max_fuse = max_fuse
        p = ctypes.c_void_p()
        check(load().swem_server_create(ctypes.byref(c), ctypes.byref(p)))
        self._s = p.value
        self._shapes: dict[int, tuple] = {}

    def submit(self, geom: WemGeometry, vel: np.ndarray, wav: np.ndarray, d: np.ndarray, sx, sy=None,
               pspi_mode: int = _lib.PSPI_COMPAT, precision: int = _lib.PREC_KZ64) -> int:
        """Queues the migration of d [nshot, nm, nt]; returns the request id.  Inputs are copied."""
        f = lambda a: np.ascontiguousarray(a, dtype=np.float32)  # noqa: E731
        vel, wav, d = f(vel), f(wav), f(d)
        sx = f(np.atleast_1d(sx))
        sy = f(np.zeros_like(sx) if sy is None else np.atleast_1d(sy))
        if wav.size < geom.nt:
            wav = np.concatenate([wav, np.zeros(geom.nt - wav.size, np.float32)])
        fp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))  # noqa: E731
        r = _lib.Request(geom.c(), pspi_mode, precision, fp(vel), fp(wav), fp(d), fp(sx), fp(sy), sx.size)
        rid = ctypes.c_longlong()
        check(load().swem_batch_submit(self._s, ctypes.byref(r), ctypes.byref(rid)))
        self._shapes[rid.value] = (geom.nm, geom.nz)
        return rid.value

    def poll(self, rid: int) -> dict:
        st = _lib.RequestStatus()
        check(load().swem_batch_poll(self._s, rid, ctypes.byref(st)))
        return self._status(st)

    def wait(self, rid: int):
        """Blocks until request rid is done; returns (image [nm, nz], status dict)."""
        nm, nz = self._shapes.pop(rid)
        m = np.empty((nm, nz), np.float32)
        st = _lib.RequestStatus()
        check(load().swem_batch_wait(self._s, rid, m.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                     ctypes.byref(st)))
        return m, self._status(st)

    def stats(self) -> dict:
        st = _lib.ServerStats()
        check(load().swem_server_get_stats(self._s, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in _lib.ServerStats._fields_}

    @staticmethod
    def _status(st) -> dict:
        return {"state": st.state, "error": st.error, "submit_ms": st.submit_ms, "admit_ms": st.admit_ms,
                "done_ms": st.done_ms, "message": st.message.decode(errors="replace")}

    def close(self):
        if getattr(self, "_s", None):
            load().swem_server_destroy(self._s)
            self._s = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
