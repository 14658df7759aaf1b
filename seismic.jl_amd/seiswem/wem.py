"""Shot-profile wave-equation migration / demigration on MI355X (host-side API).

Mirrors the reference surface:
  * ``wem(...)``              -- wem() of src/Imaging/c/wem.c:3-11 (one shot, host arrays)
  * ``ShotProfileWEM(...)``   -- src/Imaging/ShotProfileWEM.jl:39 in memory (nangx = nangy = 1)
  * ``Plan``                  -- device-resident tables for repeated calls (batched shots)
Array layouts are the reference's: d [nm, nt], m [nm, nz], vel [nm, nz], trace ix = imx*nmy + imy.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import check, load


@dataclass
class WemGeometry:
    """Scalar arguments of wem() with the Julia driver's defaults (ShotProfileWEM.jl:39)."""

    nt: int
    dt: float
    nmx: int
    dmx: float
    nz: int
    dz: float
    nmy: int = 1
    dmy: float = 0.0
    ot: float = 0.0
    omx: float = 0.0
    omy: float = 0.0
    oz: float = 0.0
    gz: float = 0.0
    sz: float = 0.0
    nref: int = 5
    fmin: float = 0.0
    fmax: float = 80.0
    padt: int = 2
    padx: int = 2
    pspi: bool = True

    @property
    def nm(self) -> int:
        return self.nmx * self.nmy

    def c(self) -> _lib.Geometry:
        dmy = self.dmy if self.dmy > 0 else self.dmx  # shotwem.c:68 sets dy = dx
        return _lib.Geometry(self.nt, self.ot, self.dt, self.nmx, self.omx, self.dmx, self.nmy, self.omy, dmy,
                             self.nz, self.oz, self.dz, self.gz, self.sz, self.nref, self.fmin, self.fmax,
                             self.padt, self.padx, int(bool(self.pspi)))

    def sampling(self) -> dict:
        out = (ctypes.c_int * 8)()
        g = self.c()
        check(load().swem_sampling(ctypes.byref(g), out))
        keys = ("ntfft", "nw", "nkx", "nky", "ifmin", "ifmax", "lmx", "lmy")
        return dict(zip(keys, list(out)))


def _fp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class DeviceArray:
    """A float32 buffer in device memory owned by the library allocator."""

    def __init__(self, shape, data: np.ndarray | None = None):
        self.shape = tuple(int(s) for s in shape)
        self.nbytes = int(np.prod(self.shape)) * 4
        p = ctypes.c_void_p()
        check(load().swem_malloc(ctypes.byref(p), max(self.nbytes, 4)))
        self.ptr = p.value
        if data is not None:
            self.upload(data)

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float32)
        if a.nbytes != self.nbytes:
            raise ValueError(f"upload of {a.shape} into {self.shape}")
        check(load().swem_memcpy_htod(self.ptr, a.ctypes.data, self.nbytes))

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, np.float32)
        check(load().swem_memcpy_dtoh(out.ctypes.data, self.ptr, self.nbytes))
        return out

    def zero(self):
        check(load().swem_memset(self.ptr, 0, self.nbytes))

    def free(self):
        if self.ptr:
            load().swem_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Plan:
    """Velocity-derived tables and the source spectrum resident on the device (swem_plan_*)."""

    def __init__(self, geom: WemGeometry, vel: np.ndarray, wav: np.ndarray, *, pspi_mode: int = _lib.PSPI_COMPAT,
                 precision: int = _lib.PREC_KZ64, freqs_per_block: int = 0, scratch_bytes: int = 0, device: int = 0,
                 variant: int = _lib.VARIANT_AUTO, table_bytes: int = 0):
        self.geom = geom
        vel = np.ascontiguousarray(vel, dtype=np.float32).reshape(geom.nm, geom.nz)
        wav = np.ascontiguousarray(wav, dtype=np.float32).reshape(-1)
        if wav.size < geom.nt:
            wav = np.concatenate([wav, np.zeros(geom.nt - wav.size, np.float32)])
        wav = np.ascontiguousarray(wav[: geom.nt])
        g = geom.c()
        o = _lib.options(pspi_mode, precision, device, freqs_per_block, False, scratch_bytes, variant, table_bytes)
        p = ctypes.c_void_p()
        check(load().swem_plan_create(ctypes.byref(g), _fp(vel), _fp(wav), ctypes.byref(o), ctypes.byref(p)))
        self._p = p.value

    def close(self):
        if getattr(self, "_p", None):
            load().swem_plan_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _shots(sx, sy):
        sx = np.ascontiguousarray(np.atleast_1d(sx), dtype=np.float32)
        sy = np.zeros_like(sx) if sy is None else np.ascontiguousarray(np.atleast_1d(sy), dtype=np.float32)
        return sx, sy

    def migrate(self, d: DeviceArray, sx, sy=None, m: DeviceArray | None = None, accumulate: bool = False,
                stream=None) -> DeviceArray:
        sx, sy = self._shots(sx, sy)
        if m is None:
            m = DeviceArray((self.geom.nm, self.geom.nz))
        check(load().swem_plan_migrate(self._p, d.ptr, _fp(sx), _fp(sy), sx.size, m.ptr, int(accumulate), stream))
        return m

    def migrate_each(self, d: DeviceArray, sx, sy=None, m: DeviceArray | None = None, stream=None) -> DeviceArray:
        sx, sy = self._shots(sx, sy)
        if m is None:
            m = DeviceArray((sx.size, self.geom.nm, self.geom.nz))
        check(load().swem_plan_migrate_each(self._p, d.ptr, _fp(sx), _fp(sy), sx.size, m.ptr, stream))
        return m

    def model(self, m: DeviceArray, sx, sy=None, d: DeviceArray | None = None, stream=None) -> DeviceArray:
        sx, sy = self._shots(sx, sy)
        if d is None:
            d = DeviceArray((sx.size, self.geom.nm, self.geom.nt))
        check(load().swem_plan_model(self._p, m.ptr, _fp(sx), _fp(sy), sx.size, d.ptr, stream))
        return d

    def timing(self) -> dict:
        t = _lib.Timing()
        check(load().swem_plan_timing(self._p, ctypes.byref(t)))
        return {k: getattr(t, k) for k, _ in _lib.Timing._fields_}

    def autotune(self, nshot: int = 1) -> dict:
        """Times every 2-D migration kernel variant that fits this plan on nshot pseudo-random
        records and keeps the fastest (swem_plan_autotune).  Returns {"best": variant, "ms": {variant: ms}}."""
        best = ctypes.c_int(-1)
        times = (ctypes.c_float * (_lib.VARIANT_PACKED + 1))()
        check(load().swem_plan_autotune(self._p, int(nshot), ctypes.byref(best), times))
        return {"best": best.value, "ms": {v: float(times[v]) for v in range(len(times)) if times[v] > 0}}

    def set_graphs(self, enable: bool):
        """3-D migration: replay each shot's launches as one captured graph (default) or launch them
        one by one (swem_plan_set_graphs)."""
        check(load().swem_plan_set_graphs(self._p, int(bool(enable))))

    def autotune_shapes(self, nshot: int = 1) -> dict:
        """swem_plan_autotune_ex: times the plan's candidate launch shapes (2-D: kernel variant x
        frequencies per workgroup; 3-D: frequencies per y-pass workgroup x step state per launch)
        and keeps the fastest.  Returns the choice and every candidate's march time."""
        r = _lib.TuneResult()
        check(load().swem_plan_autotune_ex(self._p, int(nshot), ctypes.byref(r)))
        cands = [{"variant": r.cand_variant[i], "freqs_per_block": r.cand_fpb[i],
                  "chunk_bytes": r.cand_chunk_bytes[i], "ms": r.cand_ms[i]} for i in range(r.ncand)]
        return {"variant": r.variant, "freqs_per_block": r.freqs_per_block, "chunk_bytes": r.chunk_bytes,
                "best_ms": r.best_ms, "candidates": cands}

    def bytes(self):
        a, b = ctypes.c_size_t(), ctypes.c_size_t()
        check(load().swem_plan_bytes(self._p, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value


def synchronize():
    check(load().swem_synchronize())


def copy_device(dst: DeviceArray, src: DeviceArray):
    """Device-to-device copy of a whole buffer."""
    if dst.nbytes != src.nbytes:
        raise ValueError("copy between buffers of different sizes")
    check(load().swem_memcpy_dtod(dst.ptr, src.ptr, src.nbytes))


def wem(d, m, wav, nt, ot, dt, nmx, omx, dmx, nmy, omy, dmy, sx, sy, nz, oz, dz, gz, sz, vel, nref, fmin, fmax,
        padt, padx, adj, pspi, verbose=False):
    """wem() (wem.c:3-11) on host arrays: adj fills m [nm, nz] from d [nm, nt], else d from m."""
    nm = nmx * nmy
    d_a = np.ascontiguousarray(d, dtype=np.float32) if d is not None else np.zeros((nm, nt), np.float32)
    m_a = np.ascontiguousarray(m, dtype=np.float32) if m is not None else np.zeros((nm, nz), np.float32)
    vel = np.ascontiguousarray(vel, dtype=np.float32)
    wav = np.ascontiguousarray(wav, dtype=np.float32)
    check(load().swem_wem(_fp(d_a), _fp(m_a), _fp(wav), nt, ot, dt, nmx, omx, dmx, nmy, omy, dmy, sx, sy, nz, oz, dz,
                          gz, sz, _fp(vel), nref, fmin, fmax, padt, padx, int(adj), int(pspi), int(verbose)))
    return m_a if adj else d_a


def ShotProfileWEM(x: np.ndarray, adj: bool = True, *, vel: np.ndarray, wav: np.ndarray, geom: WemGeometry,
                   sx=(0.0,), sy=None, pspi_mode: int = _lib.PSPI_COMPAT, precision: int = _lib.PREC_KZ64,
                   freqs_per_block: int = 0, variant: int = _lib.VARIANT_AUTO) -> np.ndarray:
    """In-memory ShotProfileWEM (ShotProfileWEM.jl:39-310) with nangx = nangy = 1.

    adj=True:  x = d [nshot, nm, nt]  ->  stacked image m [nm, nz] (sum over shots, :140-201)
    adj=False: x = m [nm, nz]          ->  d [nshot, nm, nt] (one record per shot, :288-308)
    """
    sx = np.ascontiguousarray(np.atleast_1d(sx), dtype=np.float32)
    sy = np.zeros_like(sx) if sy is None else np.ascontiguousarray(np.atleast_1d(sy), dtype=np.float32)
    ns = sx.size
    g = geom.c()
    o = _lib.options(pspi_mode, precision, 0, freqs_per_block, variant=variant)
    vel = np.ascontiguousarray(vel, dtype=np.float32)
    wav = np.ascontiguousarray(wav, dtype=np.float32)
    if adj:
        d = np.ascontiguousarray(x, dtype=np.float32).reshape(ns, geom.nm, geom.nt)
        m = np.zeros((geom.nm, geom.nz), np.float32)
    else:
        m = np.ascontiguousarray(x, dtype=np.float32).reshape(geom.nm, geom.nz)
        d = np.zeros((ns, geom.nm, geom.nt), np.float32)
    check(load().swem_shots(ctypes.byref(g), _fp(vel), _fp(wav), _fp(sx), _fp(sy), ns, _fp(d), _fp(m), int(adj),
                            ctypes.byref(o)))
    return m if adj else d
