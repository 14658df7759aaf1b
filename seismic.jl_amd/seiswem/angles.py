"""Angle-domain shot-profile imaging.

  compute_angles         wem.c:566-992 (compute_angles.c:12-118 without the file I/O): source
                         incidence angles of one shot on the device (2-D, and 3-D with nmx,
                         nmy up to 512 or powers of two up to 1024)
  angle_bin / angle_unbin   the bilinear angle binning of ShotProfileWEM.jl:162-188 (migration:
                         a shot image spread into angle gathers) and its transpose
                         ShotProfileWEM.jl:246-275 (modelling: angle gathers collapsed to a shot
                         image with the same weights)
  ShotProfileWEMAngles   ShotProfileWEM.jl:39-310 with nangx / nangy > 1 on arrays

Image layout with angles: [nm, nangx * nangy, nz] (trace imx * ny * nang + imy * nang + iang of
the reference's file, iang = iangx * nangy + iangy).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check, load
from .wem import DeviceArray, Plan, WemGeometry, _fp


def compute_angles(vel, wav, geom: WemGeometry, sx: float, sy: float = 0.0, verbose: bool = False, **opts):
    """Source incidence angles (degrees) angx, angy [nm, nz] of the shot at (sx, sy): the
    reference's compute_angles with its own sampling (ntfft = nt, nkx = nmx, nky = nmy; padt,
    padx, pspi and nref of `geom` are not used).  angy is zero in 2-D."""
    g = geom
    nm = g.nmx * g.nmy
    angx = np.zeros((nm, g.nz), np.float32)
    angy = np.zeros((nm, g.nz), np.float32)
    vel = np.ascontiguousarray(vel, dtype=np.float32)
    wav = np.ascontiguousarray(wav, dtype=np.float32).ravel()[: g.nt]
    if opts:  # explicit options (device, scratch_bytes, table_bytes ...): swem_compute_angles_ex
        o = _lib.options(verbose=verbose, **opts)
        geo = g.c()
        check(load().swem_compute_angles_ex(_fp(angx), _fp(angy), _fp(wav), ctypes.byref(geo), float(sx), float(sy),
                                            _fp(vel), ctypes.byref(o)))
        return angx, angy
    dmy = g.dmy if g.dmy > 0 else g.dmx
    check(load().swem_compute_angles(_fp(angx), _fp(angy), _fp(wav), g.nt, g.ot, g.dt, g.nmx, g.omx, g.dmx, g.nmy,
                                     g.omy, dmy, float(sx), float(sy), g.nz, g.oz, g.dz, g.sz, _fp(vel), g.fmin,
                                     g.fmax, int(verbose)))
    return angx, angy


def _weights(ax, ay, nangx, oangx, dangx, nangy, oangy, dangy):
    """Bins and bilinear weights of ShotProfileWEM.jl:165-177 (0-based bins)."""
    ax = np.asarray(ax, np.float64)
    ay = np.asarray(ay, np.float64)
    ix = np.floor((ax - oangx) / dangx).astype(np.int64)
    iy = np.floor((ay - oangy) / dangy).astype(np.int64)
    x1 = ix * dangx + oangx
    x2 = x1 + dangx
    y1 = iy * dangy + oangy
    y2 = y1 + dangy
    den = (x2 - x1) * (y2 - y1)
    w1 = (x2 - ax) * (y2 - ay) / den
    w2 = (ax - x1) * (y2 - ay) / den
    w3 = (x2 - ax) * (ay - y1) / den
    w4 = (ax - x1) * (ay - y1) / den
    return ix, iy, (w1, w2, w3, w4)


def _taps(ix, iy, w, nangx, nangy):
    """(bin, weight, mask) of the four corners as the reference applies them: the point counts
    only when its lower-left bin is inside; the +x corner needs iangx < nangx; the +y corners
    apply when nangy > 1 (the reference's test on the last corner names undefined variables,
    ShotProfileWEM.jl:183; the forward's test iangx < nangx, :262, is used for both), and a
    +y corner past the last bin is dropped."""
    nang = nangx * nangy
    inside = (ix >= 0) & (ix < nangx) & (iy >= 0) & (iy < nangy)
    base = ix * nangy + iy
    taps = [(base, w[0], inside), (base + nangy, w[1], inside & (ix < nangx - 1))]
    if nangy > 1:
        taps.append((base + 1, w[2], inside & (base + 1 < nang)))
        taps.append((base + nangy + 1, w[3], inside & (ix < nangx - 1) & (base + nangy + 1 < nang)))
    return taps


def angle_bin(m_shot, angx, angy, nangx, oangx, dangx, nangy=1, oangy=0.0, dangy=1.0, out=None):
    """ShotProfileWEM.jl:162-188: add the shot image m_shot [nm, nz] into the angle gathers
    out [nm, nangx * nangy, nz] with bilinear weights of the shot's angles [nm, nz]."""
    m_shot = np.asarray(m_shot, np.float32)
    nm, nz = m_shot.shape
    nang = nangx * nangy
    out = np.zeros((nm, nang, nz), np.float32) if out is None else out
    ix, iy, w = _weights(angx, angy, nangx, oangx, dangx, nangy, oangy, dangy)
    rows = np.arange(nm)[:, None].repeat(nz, 1)
    cols = np.arange(nz)[None, :].repeat(nm, 0)
    for b, wt, mask in _taps(ix, iy, w, nangx, nangy):
        v = (wt * m_shot)[mask]
        np.add.at(out, (rows[mask], b[mask], cols[mask]), v.astype(np.float32))
    return out


def angle_unbin(m_ang, angx, angy, nangx, oangx, dangx, nangy=1, oangy=0.0, dangy=1.0):
    """ShotProfileWEM.jl:246-275: the shot image [nm, nz] gathered from the angle gathers
    m_ang [nm, nangx * nangy, nz] with the same weights (the transpose of angle_bin)."""
    m_ang = np.asarray(m_ang, np.float32)
    nm, nang, nz = m_ang.shape
    ix, iy, w = _weights(angx, angy, nangx, oangx, dangx, nangy, oangy, dangy)
    out = np.zeros((nm, nz), np.float64)
    rows = np.arange(nm)[:, None].repeat(nz, 1)
    cols = np.arange(nz)[None, :].repeat(nm, 0)
    for b, wt, mask in _taps(ix, iy, w, nangx, nangy):
        bb = np.where(mask, b, 0)
        out += np.where(mask, wt * m_ang[rows, bb, cols], 0.0)
    return out.astype(np.float32)


def ShotProfileWEMAngles(x, adj: bool = True, *, vel, wav, geom: WemGeometry, sx, sy=None, nangx=1, oangx=0.0,
                         dangx=1.0, nangy=1, oangy=0.0, dangy=1.0, angx=None, angy=None, **plan_kw):
    """ShotProfileWEM.jl:39-310 on arrays with angle gathers.  adj: records x [nshot, nm, nt] ->
    angle-domain image [nm, nangx * nangy, nz] (each shot migrated on the device, spread by its
    incidence angles); else the image -> records [nshot, nm, nt].  The angles of each shot
    ([nshot, nm, nz], the reference's angx / angy files) are computed on the device when not
    given; with nangx = nangy = 1 the angles are zero (ShotProfileWEM.jl:151-154)."""
    g = geom
    sx = np.atleast_1d(np.asarray(sx, np.float32))
    ns = sx.size
    syv = np.zeros(ns, np.float32) if sy is None else np.atleast_1d(np.asarray(sy, np.float32))
    use = nangx != 1 or nangy != 1
    if angx is None or angy is None:
        ax = np.zeros((ns, g.nm, g.nz), np.float32)
        ay = np.zeros((ns, g.nm, g.nz), np.float32)
        if use:
            for i in range(ns):
                ax[i], ay[i] = compute_angles(vel, wav, g, float(sx[i]), float(syv[i]))
        angx, angy = ax, ay
    plan = Plan(g, vel, wav, **plan_kw)
    try:
        if adj:
            d = DeviceArray((ns, g.nm, g.nt), np.asarray(x, np.float32))
            each = plan.migrate_each(d, sx, None if sy is None else syv).download()
            out = np.zeros((g.nm, nangx * nangy, g.nz), np.float32)
            for i in range(ns):
                angle_bin(each[i], angx[i], angy[i], nangx, oangx, dangx, nangy, oangy, dangy, out=out)
            return out
        m_ang = np.asarray(x, np.float32).reshape(g.nm, nangx * nangy, g.nz)
        d = np.zeros((ns, g.nm, g.nt), np.float32)
        for i in range(ns):
            m_shot = angle_unbin(m_ang, angx[i], angy[i], nangx, oangx, dangx, nangy, oangy, dangy)
            d[i] = plan.model(DeviceArray(m_shot.shape, m_shot), sx[i:i + 1],
                              None if sy is None else syv[i:i + 1]).download()[0]
        return d
    finally:
        plan.close()
