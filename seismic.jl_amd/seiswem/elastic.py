"""Elastic shot-profile wave-equation migration and modelling (2-D).

  ewem             ewem.c:3-15 on host arrays (swem_ewem): components ux, uy, uz [nm, nt] <->
                   images mpp, mps1, mps2 [nm, nz]
  ShotProfileEWEM  ShotProfileEWEM.jl on arrays with nangx = nangy = 1: several shots over one
                   pair of velocity grids, images stacked (migration) or records per shot
                   (modelling)
"""
from __future__ import annotations

import numpy as np

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
    """ShotProfileEWEM (ShotProfileEWEM.jl) on arrays.  adj: x = (ux, uy, uz), each
    [nshot, nm, nt] -> (mpp, mps1, mps2) [nm, nz] summed over shots; else x = (mpp, mps1, mps2)
    -> (ux, uy, uz) [nshot, nm, nt]."""
    g = geom
    sx = np.atleast_1d(np.asarray(sx, np.float32))
    sy = np.zeros_like(sx) if sy is None else np.atleast_1d(np.asarray(sy, np.float32))
    dmy = g.dmy if g.dmy > 0 else g.dmx
    common = (g.nt, g.ot, g.dt, g.nmx, g.omx, g.dmx, g.nmy, g.omy, dmy)
    if adj:
        out = [np.zeros((g.nm, g.nz), np.float32) for _ in range(3)]
        for i in range(sx.size):
            m = ewem(x[0][i], x[1][i], x[2][i], None, None, None, wav, *common, float(sx[i]), float(sy[i]), g.nz, g.oz,
                     g.dz, g.gz, g.sz, vp, vs, g.fmin, g.fmax, g.padt, g.padx, True)
            for k in range(3):
                out[k] += m[k]
        return tuple(out)
    out = [np.zeros((sx.size, g.nm, g.nt), np.float32) for _ in range(3)]
    for i in range(sx.size):
        d = ewem(None, None, None, x[0], x[1], x[2], wav, *common, float(sx[i]), float(sy[i]), g.nz, g.oz, g.dz, g.gz,
                 g.sz, vp, vs, g.fmin, g.fmax, g.padt, g.padx, False)
        for k in range(3):
            out[k][i] = d[k]
    return tuple(out)
