"""Elastic shot-profile wave-equation migration and modelling (2-D).

  ewem             ewem.c:3-15 on host arrays (swem_ewem): components ux, uy, uz [nm, nt] <->
                   images mpp, mps1, mps2 [nm, nz]
  ShotProfileEWEM  ShotProfileEWEM.jl on arrays with nangx = nangy = 1: several shots over one
                   pair of velocity grids, images stacked (migration) or records per shot
                   (modelling)
"""
from __future__ import annotations

import numpy as np

import ctypes

from . import _lib
from ._lib import check, load
from .wem import WemGeometry, _fp


def ewem(ux, uy, uz, mpp, mps1, mps2, wav, nt, ot, dt, nmx, omx, dmx, nmy, omy, dmy, sx, sy, nz, oz, dz, gz, sz,
         vp, vs, fmin, fmax, padt, padx, adj, verbose=False):
    """ewem() (ewem.c:3-15): adj fills (mpp, mps1, mps2) [nm, nz] from (ux, uy, uz) [nm, nt];
    else the components from the images.  Returns the three outputs."""
    nm = nmx * nmy

    def arr(a, shape):
        return np.ascontiguousarray(a, dtype=np.float32) if a is not None else np.zeros(shape, np.float32)

    d = [arr(a, (nm, nt)) for a in (ux, uy, uz)]
    m = [arr(a, (nm, nz)) for a in (mpp, mps1, mps2)]
    vp = np.ascontiguousarray(vp, dtype=np.float32)
    vs = np.ascontiguousarray(vs, dtype=np.float32)
    wav = np.ascontiguousarray(wav, dtype=np.float32).ravel()
    check(load().swem_ewem(*[_fp(a) for a in d + m], _fp(wav), nt, ot, dt, nmx, omx, dmx, nmy, omy, dmy, sx, sy, nz,
                           oz, dz, gz, sz, _fp(vp), _fp(vs), fmin, fmax, padt, padx, int(adj), int(verbose)))
    return tuple(m) if adj else tuple(d)


def ShotProfileEWEM(x, adj: bool = True, *, vp, vs, wav, geom: WemGeometry, sx=(0.0,), sy=None):
    """ShotProfileEWEM (ShotProfileEWEM.jl) on arrays, one device call for all shots
    (swem_eshots).  adj: x = (ux, uy, uz), each [nshot, nm, nt] -> (mpp, mps1, mps2) [nm, nz]
    summed over shots; else x = (mpp, mps1, mps2) -> (ux, uy, uz) [nshot, nm, nt]."""
    g = geom
    sx = np.ascontiguousarray(np.atleast_1d(sx), dtype=np.float32)
    ns = sx.size
    syv = None if sy is None else np.ascontiguousarray(np.atleast_1d(sy), dtype=np.float32)
    if adj:
        d = [np.ascontiguousarray(a, dtype=np.float32).reshape(ns, g.nm, g.nt) for a in x]
        m = [np.zeros((g.nm, g.nz), np.float32) for _ in range(3)]
    else:
        m = [np.ascontiguousarray(a, dtype=np.float32).reshape(g.nm, g.nz) for a in x]
        d = [np.zeros((ns, g.nm, g.nt), np.float32) for _ in range(3)]
    geo = g.c()
    opts = _lib.options()
    check(load().swem_eshots(ctypes.byref(geo), _fp(np.ascontiguousarray(vp, np.float32)),
                             _fp(np.ascontiguousarray(vs, np.float32)),
                             _fp(np.ascontiguousarray(wav, np.float32).ravel()), _fp(sx),
                             None if syv is None else _fp(syv), ns, *[_fp(a) for a in d + m], int(adj), ctypes.byref(opts)))
    return tuple(m) if adj else tuple(d)
