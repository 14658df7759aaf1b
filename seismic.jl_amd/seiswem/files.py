"""File-based ShotProfileWEM (src/Imaging/ShotProfileWEM.jl:39-310) for nangx = nangy = 1.

The reference windows every shot out of the data file by its (sx, sy) headers, runs one shotwem
process per shot through files and stacks the per-shot images into m at trace position
imx * ny + imy (migration), or writes the image per shot and concatenates the modelled shots
(demigration).  Here the shots of one call share one device plan (velocity tables built once)
and run in one batch; the files read and written, their extents and headers are the
reference's:
  migration  m: extent (nz, nx, ny, 1, 1), labels Depth/mx/my/angx/angy; headers imx/imy/mx/my
             of the velocity grid (ShotProfileWEM.jl:112-126) with tracenum the 1-based trace
             number that SeisWrite stores (SeisWrite.jl:36-39); the image of each shot is added
             at trace imx * ny + imy, imx = trunc((gx - ox)/dx) of its data traces
             (shotwem.c:113-117, ShotProfileWEM.jl:145-150).
  modelling  d: extent (nt, nx, ny, nsx, nsy), labels Time/gx/gy/sx/sy; m is read by position,
             trace (imx + min_imx) * ny + imy + min_imy (ShotProfileWEM.jl:243-247); per shot the
             traces of the velocity grid with gx = mx, hx = mx - sx (shotwem.c:86-98), shots
             concatenated with tracenum running over the whole file.
With one angle bin the bilinear angle weights of ShotProfileWEM.jl:162-188 are 1 and 0."""
from __future__ import annotations

import numpy as np

from . import seis
from .wem import DeviceArray, Plan, WemGeometry


def _scale(d, i):
    """Float32 d times an integer index, rounded once as Julia's Float32 * Int."""
    return np.float32(d) * np.asarray(i).astype(np.float32)


def _grid(vel_name):
    v, h, fh = seis.read(vel_name)
    min_imx, max_imx = int(h[0]["imx"]), int(h[-1]["imx"])
    min_imy, max_imy = int(h[0]["imy"]), int(h[-1]["imy"])
    nx, ny = max_imx - min_imx + 1, max_imy - min_imy + 1
    ox = float(h[0]["mx"])
    dx = float(h[1]["mx"] - h[0]["mx"]) if h.size > 1 else float(fh.d2)
    oy = float(h[0]["my"])
    dy = float(h[nx]["my"] - oy) if ny > 1 else dx
    # The Julia driver takes nz/dz/oz from the first trace header (ShotProfileWEM.jl:54-56), shotwem
    # from the file header and sets dy = dx (shotwem.c:47, 63-68): both are kept.
    return dict(v=v, h=h, nx=nx, ny=ny, ox=ox, dx=dx, oy=oy, dy=dy, min_imx=min_imx, min_imy=min_imy,
                nz=int(h[0]["n1"]), dz=float(h[0]["d1"]), oz=float(h[0]["o1"]), fnz=fh.n1, fdz=fh.d1, foz=fh.o1)


def ShotProfileWEMFiles(m: str, d: str, adj: bool = True, *, pspi=True, nref=5, vel="vel", wav="wav", sz=0.0,
                        gz=0.0, fmin=0.0, fmax=80.0, padt=2, padx=2, sx=(0.0,), sy=(0.0,), pspi_mode: int = 0):
    """ShotProfileWEM(m, d, adj; ...) on .seis files (nangx = nangy = 1)."""
    gr = _grid(vel)
    w, hw, fw = seis.read(wav)
    nt, dt, ot = int(hw[0]["n1"]), float(hw[0]["d1"]), float(hw[0]["o1"])
    nx, ny, nz = gr["nx"], gr["ny"], gr["nz"]
    if (nt, nz) != (fw.n1, gr["fnz"]) or nx * ny != gr["v"].shape[0]:
        raise ValueError("wavelet / velocity headers disagree with their file headers")
    nm = nx * ny
    # SeisWindowHeaders compares float32 header words with float32 limits (SeisWindowHeaders.jl:33-34).
    # The shot list is sx with sy[ishot] (ShotProfileWEM.jl:41, 80); the extent counts length(sy).
    sx = np.atleast_1d(np.asarray(sx, np.float32))
    sy_all = np.atleast_1d(np.asarray(sy, np.float32))
    if sy_all.size < sx.size:
        raise ValueError("sy needs an entry for every shot in sx")
    sy = sy_all[: sx.size]
    geom = WemGeometry(nt=nt, ot=fw.o1, dt=fw.d1, nmx=nx, omx=gr["ox"], dmx=gr["dx"], nmy=ny, omy=gr["oy"],
                       dmy=gr["dx"], nz=nz, oz=gr["foz"], dz=gr["fdz"], gz=gz, sz=sz, nref=nref, fmin=fmin,
                       fmax=fmax, padt=padt, padx=padx, pspi=pspi)
    plan = Plan(geom, gr["v"], w[0], pspi_mode=pspi_mode)
    if adj:
        dd, hd, fd = seis.read(d)
        shots = []
        for s_x, s_y in zip(sx, sy):
            sel = (hd["sx"] == s_x) & (hd["sy"] == s_y)
            if sel.sum() != nm:
                raise ValueError(f"shot ({s_x}, {s_y}): {sel.sum()} traces, the velocity grid has {nm}")
            shots.append(np.nonzero(sel)[0])
        dev = DeviceArray((sx.size, nm, nt), np.stack([dd[i] for i in shots]))
        each = plan.migrate_each(dev, sx, sy).download()
        image = np.zeros((nm, nz), np.float32)
        for k, (idx, s_x, s_y) in enumerate(zip(shots, sx, sy)):
            imx = np.trunc((hd["gx"][idx] - gr["ox"]) / gr["dx"]).astype(int)
            imy = np.trunc((hd["gy"][idx] - gr["oy"]) / gr["dx"]).astype(int)
            pos = imx * ny + imy
            if pos.min() < 0 or pos.max() >= nm:
                raise ValueError(f"shot ({s_x}, {s_y}): receiver outside the velocity grid")
            np.add.at(image, pos, each[k])
        hm = seis.headers(nm)
        j = np.arange(nm)
        imx, imy = j // ny + gr["min_imx"], j % ny + gr["min_imy"]
        hm["tracenum"] = j + 1
        hm["o1"], hm["n1"], hm["d1"] = 0.0, nz, gr["dz"]
        hm["imx"], hm["imy"] = imx, imy
        hm["mx"], hm["my"] = _scale(gr["dx"], imx), _scale(gr["dy"], imy)
        fh = seis.extent((nz, nx, ny, 1, 1), (gr["oz"], gr["ox"], gr["oy"], 0.0, 0.0),
                         (gr["dz"], gr["dx"], gr["dy"], 1.0, 1.0), ("Depth", "mx", "my", "angx", "angy"),
                         ("m", "m", "m", "degrees", "degrees"))
        seis.write(m, image, hm, fh)
        return image
    mm, _, _ = seis.read(m)
    j = np.arange(nm)
    pos = (j // ny + gr["min_imx"]) * ny + j % ny + gr["min_imy"]
    if mm.shape[1] != nz or pos.max() >= mm.shape[0]:
        raise ValueError(f"{m}: needs {pos.max() + 1} traces of {nz} samples")
    image = np.ascontiguousarray(mm[pos])
    rec = plan.model(DeviceArray((nm, nz), image), sx, sy).download()
    hd = seis.headers(sx.size * nm)
    for k in range(sx.size):
        h = hd[k * nm:(k + 1) * nm]
        imx, imy = j // ny + gr["min_imx"], j % ny + gr["min_imy"]
        h["tracenum"] = k * nm + j + 1
        h["o1"], h["n1"], h["d1"] = fw.o1, nt, fw.d1
        h["imx"], h["imy"] = imx, imy
        h["mx"], h["my"] = _scale(gr["dx"], imx), _scale(gr["dy"], imy)
        h["sx"], h["sy"] = sx[k], sy[k]
        h["gx"], h["gy"] = h["mx"], h["my"]
        h["hx"], h["hy"] = h["mx"] - sx[k], h["my"] - sy[k]
    osx, osy = float(sx[0]), float(sy_all[0])
    dsx = float(sx[1] - sx[0]) if sx.size > 1 else 1.0
    dsy = float(sy_all[1] - sy_all[0]) if sy_all.size > 1 else 1.0
    fh = seis.extent((nt, nx, ny, sx.size, sy_all.size), (ot, gr["ox"], gr["oy"], osx, osy),
                     (dt, gr["dx"], gr["dy"], dsx or 1.0, dsy or 1.0), ("Time", "gx", "gy", "sx", "sy"),
                     ("s", "m", "m", "m", "m"))
    seis.write(d, rec.reshape(sx.size * nm, nt), hd, fh)
    return rec
