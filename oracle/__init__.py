"""CPU oracle for shot-profile wave-equation migration -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import this package, and only as the checker.  The product path (``seismic.jl_amd``) never
imports it and fails loudly when its HIP library is missing.

The oracle is a C restatement of the reference path (``src/Imaging/c/wem.c:3-564``) built from
``oracle/wem_oracle.c`` into ``oracle/_build/liboracle.so``.  ``*64`` functions evaluate every
continuous quantity in double precision (the FP64 oracle); ``*32`` functions use the
reference's float expressions (naive FP32).  Discrete decisions (sampling, source index,
reference-velocity indices, the PSPI branch test, taper widths) follow the reference's casts in
both.  Parity status of each function is recorded in tests/test_oracle_*.py.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None


class Geom(ctypes.Structure):
    """orc_geom (oracle/wem_oracle.h) -- the parameters of one wem() call (wem.c:3-11)."""

    _fields_ = [
        ("nt", ctypes.c_int), ("ot", ctypes.c_float), ("dt", ctypes.c_float),
        ("nmx", ctypes.c_int), ("omx", ctypes.c_float), ("dmx", ctypes.c_float),
        ("nmy", ctypes.c_int), ("omy", ctypes.c_float), ("dmy", ctypes.c_float),
        ("nz", ctypes.c_int), ("oz", ctypes.c_float), ("dz", ctypes.c_float),
        ("gz", ctypes.c_float), ("sz", ctypes.c_float),
        ("nref", ctypes.c_int), ("fmin", ctypes.c_float), ("fmax", ctypes.c_float),
        ("padt", ctypes.c_int), ("padx", ctypes.c_int),
        ("pspi", ctypes.c_int), ("pspi_mode", ctypes.c_int),
    ]


class Sampling(ctypes.Structure):
    _fields_ = [
        ("ntfft", ctypes.c_int), ("nw", ctypes.c_int), ("nkx", ctypes.c_int), ("nky", ctypes.c_int),
        ("ifmin", ctypes.c_int), ("ifmax", ctypes.c_int), ("lmx", ctypes.c_int), ("lmy", ctypes.c_int),
        ("fop_n", ctypes.c_int),
        ("dkx", ctypes.c_float), ("dky", ctypes.c_float), ("dw", ctypes.c_float),
    ]


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile (gcc + OpenMP)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []), check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.POINTER
        f32p, f64p = P(ctypes.c_float), P(ctypes.c_double)
        i32p, u8p = P(ctypes.c_int), P(ctypes.c_ubyte)
        gp = P(Geom)
        L.orc_sampling_compute.argtypes = [gp, P(Sampling)]
        L.orc_source_index.argtypes = [gp, ctypes.c_float, ctypes.c_float, i32p, i32p]
        L.orc_pspi_tables.argtypes = [gp, f32p, f32p, i32p, i32p, u8p]
        L.orc_slowness_f32.argtypes = [gp, f32p, f32p, f32p]
        L.orc_taper_f32.argtypes = [ctypes.c_int, ctypes.c_int, f32p]
        L.orc_taper_f32.restype = None
        L.orc_taper_f64.argtypes = [ctypes.c_int, ctypes.c_int, f64p]
        L.orc_taper_f64.restype = None
        for pre, rp in (("orc64_", f64p), ("orc32_", f32p)):
            getattr(L, pre + "fft").argtypes = [rp, ctypes.c_int, ctypes.c_int]
            getattr(L, pre + "fop").argtypes = [rp, rp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
            getattr(L, pre + "step").argtypes = [rp, f32p, gp, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_int]
            getattr(L, pre + "wem").argtypes = [rp, rp, f32p, f32p, gp, ctypes.c_float, ctypes.c_float, ctypes.c_int]
            getattr(L, pre + "shots").argtypes = [rp, rp, f32p, f32p, gp, f32p, f32p, ctypes.c_int, ctypes.c_int]
            getattr(L, pre + "angles").argtypes = [rp, rp, f32p, f32p, gp, ctypes.c_float, ctypes.c_float]
            getattr(L, pre + "ewem").argtypes = [rp, rp, rp, rp, f32p, f32p, f32p, gp, ctypes.c_float, ctypes.c_float,
                                                 ctypes.c_int]
            st = ctypes.c_double if pre == "orc64_" else ctypes.c_float
            getattr(L, pre + "esep_k").argtypes = [st, st, st, st, ctypes.c_int, ctypes.c_int, rp, rp, rp, rp]
            getattr(L, pre + "esep_k").restype = None
        _lib = L
    return _lib


def _ptr(a: np.ndarray, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def _check(rc: int, what: str):
    if rc != 0:
        raise ValueError(f"oracle {what} failed with code {rc}")


@dataclass
class WemParams:
    """Parameters of wem() (wem.c:3-11) with the Julia driver's defaults (ShotProfileWEM.jl:39)."""

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
    pspi_mode: int = 0

    def geom(self) -> Geom:
        dmy = self.dmy if self.dmy > 0 else self.dmx  # shotwem.c:68 forces dy = dx
        return Geom(self.nt, self.ot, self.dt, self.nmx, self.omx, self.dmx, self.nmy, self.omy, dmy,
                    self.nz, self.oz, self.dz, self.gz, self.sz, self.nref, self.fmin, self.fmax,
                    self.padt, self.padx, int(bool(self.pspi)), int(self.pspi_mode))

    @property
    def nm(self) -> int:
        return self.nmx * self.nmy


def sampling(p: WemParams) -> Sampling:
    s = Sampling()
    g = p.geom()
    _check(lib().orc_sampling_compute(ctypes.byref(g), ctypes.byref(s)), "sampling")
    return s


def source_index(p: WemParams, sx: float, sy: float = 0.0):
    g = p.geom()
    ix, iy = ctypes.c_int(), ctypes.c_int()
    rc = lib().orc_source_index(ctypes.byref(g), sx, sy, ctypes.byref(ix), ctypes.byref(iy))
    return rc, ix.value, iy.value


def pspi_tables(p: WemParams, vel: np.ndarray):
    vel = np.ascontiguousarray(vel, dtype=np.float32)
    g = p.geom()
    vref = np.zeros((p.nref, p.nz), np.float32)
    i1 = np.zeros((p.nm, p.nz), np.int32)
    i2 = np.zeros((p.nm, p.nz), np.int32)
    cut = np.zeros((p.nm, p.nz), np.uint8)
    _check(lib().orc_pspi_tables(ctypes.byref(g), _ptr(vel, ctypes.c_float), _ptr(vref, ctypes.c_float),
                                 _ptr(i1, ctypes.c_int), _ptr(i2, ctypes.c_int), _ptr(cut, ctypes.c_ubyte)), "pspi_tables")
    return vref, i1, i2, cut


def slowness_f32(p: WemParams, vel: np.ndarray):
    vel = np.ascontiguousarray(vel, dtype=np.float32)
    g = p.geom()
    po = np.zeros(p.nz, np.float32)
    pd = np.zeros((p.nm, p.nz), np.float32)
    _check(lib().orc_slowness_f32(ctypes.byref(g), _ptr(vel, ctypes.c_float), _ptr(po, ctypes.c_float),
                                  _ptr(pd, ctypes.c_float)), "slowness")
    return po, pd


def taper(n: int, l: int, f64: bool = True) -> np.ndarray:
    if f64:
        t = np.zeros(n, np.float64)
        lib().orc_taper_f64(n, l, _ptr(t, ctypes.c_double))
    else:
        t = np.zeros(n, np.float32)
        lib().orc_taper_f32(n, l, _ptr(t, ctypes.c_float))
    return t


def fft(x: np.ndarray, sign: int = -1, f64: bool = True) -> np.ndarray:
    """Unnormalised complex FFT of a 1-D array; sign -1 is FFTW_FORWARD."""
    dt = np.complex128 if f64 else np.complex64
    y = np.ascontiguousarray(x, dtype=dt).copy()
    rt, ct = (np.float64, ctypes.c_double) if f64 else (np.float32, ctypes.c_float)
    fn = lib().orc64_fft if f64 else lib().orc32_fft
    _check(fn(_ptr(y.view(rt), ct), y.size, sign), "fft")
    return y


def fop(x: np.ndarray, nw: int, nt: int, adj: bool, f64: bool = True) -> np.ndarray:
    """f_op (wem.c:480-514): adj -> r2c of a length-nt trace; not adj -> c2r to nt samples."""
    rt, ct, cdt = (np.float64, ctypes.c_double, np.complex128) if f64 else (np.float32, ctypes.c_float, np.complex64)
    fn = lib().orc64_fop if f64 else lib().orc32_fop
    if adj:
        d = np.ascontiguousarray(x, dtype=rt).copy()
        m = np.zeros(nw, cdt)
    else:
        m = np.ascontiguousarray(x, dtype=cdt).copy()
        d = np.zeros(nt, rt)
    _check(fn(_ptr(m.view(rt), ct), _ptr(d, ct), nw, nt, int(adj)), "fop")
    return m if adj else d


def step(p: WemParams, vel: np.ndarray, u: np.ndarray, iw: int, iz: int, dzs: float, adj: bool,
         f64: bool = True) -> np.ndarray:
    """One ssop/pspiop call (wem.c:289-471) on a wavefield u [nm]."""
    rt, ct, cdt = (np.float64, ctypes.c_double, np.complex128) if f64 else (np.float32, ctypes.c_float, np.complex64)
    fn = lib().orc64_step if f64 else lib().orc32_step
    vel = np.ascontiguousarray(vel, dtype=np.float32)
    y = np.ascontiguousarray(u, dtype=cdt).copy()
    g = p.geom()
    _check(fn(_ptr(y.view(rt), ct), _ptr(vel, ctypes.c_float), ctypes.byref(g), iw, iz, dzs, int(adj)), "step")
    return y


def wem(p: WemParams, vel: np.ndarray, wav: np.ndarray, sx: float, sy: float = 0.0, *, d=None, m=None,
        adj: bool = True, f64: bool = True):
    """One shot of wem() (wem.c:3-204).  adj: d [nm, nt] -> m [nm, nz]; else m -> d."""
    rt, ct = (np.float64, ctypes.c_double) if f64 else (np.float32, ctypes.c_float)
    fn = lib().orc64_wem if f64 else lib().orc32_wem
    vel = np.ascontiguousarray(vel, dtype=np.float32)
    wav = np.ascontiguousarray(wav, dtype=np.float32)
    if adj:
        d = np.ascontiguousarray(d, dtype=rt)
        m = np.zeros((p.nm, p.nz), rt)
    else:
        m = np.ascontiguousarray(m, dtype=rt)
        d = np.zeros((p.nm, p.nt), rt)
    g = p.geom()
    _check(fn(_ptr(d, ct), _ptr(m, ct), _ptr(wav, ctypes.c_float), _ptr(vel, ctypes.c_float), ctypes.byref(g),
              sx, sy, int(adj)), "wem")
    return m if adj else d


def shots(p: WemParams, vel: np.ndarray, wav: np.ndarray, sx, sy=None, *, d=None, m=None, adj: bool = True,
          f64: bool = True):
    """ShotProfileWEM with nangx=nangy=1: adj d [nshot, nm, nt] -> stacked m [nm, nz]; else m -> d."""
    rt, ct = (np.float64, ctypes.c_double) if f64 else (np.float32, ctypes.c_float)
    fn = lib().orc64_shots if f64 else lib().orc32_shots
    sx = np.ascontiguousarray(sx, dtype=np.float32)
    ns = sx.size
    sy = np.zeros(ns, np.float32) if sy is None else np.ascontiguousarray(sy, dtype=np.float32)
    vel = np.ascontiguousarray(vel, dtype=np.float32)
    wav = np.ascontiguousarray(wav, dtype=np.float32)
    if adj:
        d = np.ascontiguousarray(d, dtype=rt)
        m = np.zeros((p.nm, p.nz), rt)
    else:
        m = np.ascontiguousarray(m, dtype=rt)
        d = np.zeros((ns, p.nm, p.nt), rt)
    g = p.geom()
    _check(fn(_ptr(d, ct), _ptr(m, ct), _ptr(wav, ctypes.c_float), _ptr(vel, ctypes.c_float), ctypes.byref(g),
              _ptr(sx, ctypes.c_float), _ptr(sy, ctypes.c_float), ns, int(adj)), "shots")
    return m if adj else d


def angles(p: WemParams, vel: np.ndarray, wav: np.ndarray, sx: float, sy: float = 0.0, f64: bool = True):
    """compute_angles (wem.c:566-992) of one shot: (angx, angy) [nm, nz] in degrees.  padt, padx
    and pspi of `p` are ignored (the routine uses ntfft = nt, nkx = nmx and split-step)."""
    rt, ct = (np.float64, ctypes.c_double) if f64 else (np.float32, ctypes.c_float)
    fn = lib().orc64_angles if f64 else lib().orc32_angles
    vel = np.ascontiguousarray(vel, dtype=np.float32)
    wav = np.ascontiguousarray(wav, dtype=np.float32)
    ax = np.zeros((p.nm, p.nz), rt)
    ay = np.zeros((p.nm, p.nz), rt)
    g = p.geom()
    _check(fn(_ptr(ax, ct), _ptr(ay, ct), _ptr(wav, ctypes.c_float), _ptr(vel, ctypes.c_float), ctypes.byref(g),
              sx, sy), "angles")
    return ax, ay


def ewem(p: WemParams, vp: np.ndarray, vs: np.ndarray, wav: np.ndarray, sx: float, sy: float = 0.0, *, ux=None,
         uz=None, mpp=None, mps2=None, adj: bool = True, f64: bool = True):
    """ewem (ewem.c:3-366), 2-D.  adj: components ux, uz [nm, nt] -> images (mpp, mps2) [nm, nz]
    (mps1 of the reference is always zero); else (mpp, mps2) -> (ux, uz)."""
    rt, ct = (np.float64, ctypes.c_double) if f64 else (np.float32, ctypes.c_float)
    fn = lib().orc64_ewem if f64 else lib().orc32_ewem
    vp = np.ascontiguousarray(vp, dtype=np.float32)
    vs = np.ascontiguousarray(vs, dtype=np.float32)
    wav = np.ascontiguousarray(wav, dtype=np.float32)
    shp_d, shp_m = (p.nm, p.nt), (p.nm, p.nz)
    if adj:
        ux = np.ascontiguousarray(ux, dtype=rt)
        uz = np.ascontiguousarray(uz, dtype=rt)
        mpp, mps2 = np.zeros(shp_m, rt), np.zeros(shp_m, rt)
    else:
        mpp = np.ascontiguousarray(mpp, dtype=rt)
        mps2 = np.ascontiguousarray(mps2, dtype=rt)
        ux, uz = np.zeros(shp_d, rt), np.zeros(shp_d, rt)
    g = p.geom()
    _check(fn(_ptr(ux, ct), _ptr(uz, ct), _ptr(mpp, ct), _ptr(mps2, ct), _ptr(wav, ctypes.c_float),
              _ptr(vp, ctypes.c_float), _ptr(vs, ctypes.c_float), ctypes.byref(g), sx, sy, int(adj)), "ewem")
    return (mpp, mps2) if adj else (ux, uz)


def esep_k(kx: float, w: float, vp: float, vs: float, sep: bool, adj: bool, a1: complex, a2: complex):
    """elastic_separate_3d's 2 x 2 map at one wavenumber (ewem.c:600-630, 680-710), FP64."""
    i1 = np.array([a1.real, a1.imag])
    i2 = np.array([a2.real, a2.imag])
    o1, o2 = np.zeros(2), np.zeros(2)
    P = ctypes.POINTER(ctypes.c_double)
    lib().orc64_esep_k(kx, w, vp, vs, int(sep), int(adj), i1.ctypes.data_as(P), i2.ctypes.data_as(P),
                       o1.ctypes.data_as(P), o2.ctypes.data_as(P))
    return complex(o1[0], o1[1]), complex(o2[0], o2[1])
