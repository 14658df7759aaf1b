"""File-based ShotProfileWEM (src/Imaging/ShotProfileWEM.jl:39-310).

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
With one angle bin the bilinear angle weights of ShotProfileWEM.jl:162-188 are 1 and 0.  With
nangx * nangy > 1 each shot's angles are windowed out of the angx / angy files (the output of
ComputeAngles / compute_angles.so, traces with the shot's sx, sy) and
  migration  m: extent (nz, nx, ny, nangx, nangy); trace imx * ny * nang + imy * nang + iang,
             iang = iangx * nangy + iangy, with iang / ang / iaz / az headers
             (ShotProfileWEM.jl:108-136); each shot image is spread over the bins with the
             bilinear weights of its angles (:162-188);
  modelling  each shot's image is gathered from the bins with the same weights (:246-275) and
             modelled on its own."""
from __future__ import annotations

import numpy as np

from . import seis
from .angles import angle_bin, angle_unbin
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


def _shot_angles(angx, angy, sx, sy, nm):
    """The angle traces [nm, nz] of each shot, windowed by their sx / sy headers
    (ShotProfileWEM.jl:93-98)."""
    ax, hax, _ = seis.read(angx)
    ay, hay, _ = seis.read(angy)
    out = []
    for s_x, s_y in zip(sx, sy):
        sel_x = np.nonzero((hax["sx"] == s_x) & (hax["sy"] == s_y))[0]
        sel_y = np.nonzero((hay["sx"] == s_x) & (hay["sy"] == s_y))[0]
        if sel_x.size != nm or sel_y.size != nm:
            raise ValueError(f"shot ({s_x}, {s_y}): {sel_x.size} / {sel_y.size} angle traces, the velocity grid has {nm}")
        out.append((ax[sel_x], ay[sel_y]))
    return out


def ShotProfileWEMFiles(m: str, d: str, adj: bool = True, *, pspi=True, nref=5, vel="vel", angx="angx", angy="angy",
                        wav="wav", sz=0.0, gz=0.0, nangx=1, oangx=0.0, dangx=1.0, nangy=1, oangy=0.0, dangy=1.0,
                        fmin=0.0, fmax=80.0, padt=2, padx=2, sx=(0.0,), sy=(0.0,), pspi_mode: int = 0):
    """ShotProfileWEM(m, d, adj; ...) on .seis files."""
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
    nang = nangx * nangy
    angles = _shot_angles(angx, angy, sx, sy, nm) if nang > 1 else None
    akw = dict(nangx=nangx, oangx=oangx, dangx=dangx, nangy=nangy, oangy=oangy, dangy=dangy)
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
        image = np.zeros((nm, nang, nz), np.float32)
        for k, (idx, s_x, s_y) in enumerate(zip(shots, sx, sy)):
            imx = np.trunc((hd["gx"][idx] - gr["ox"]) / gr["dx"]).astype(int)
            imy = np.trunc((hd["gy"][idx] - gr["oy"]) / gr["dx"]).astype(int)
            pos = imx * ny + imy
            if pos.min() < 0 or pos.max() >= nm:
                raise ValueError(f"shot ({s_x}, {s_y}): receiver outside the velocity grid")
            if angles is None:
                np.add.at(image[:, 0, :], pos, each[k])
            else:
                g_k = angle_bin(each[k], angles[k][0], angles[k][1], **akw)
                np.add.at(image, pos, g_k)
        hm = seis.headers(nm * nang)
        j = np.arange(nm * nang)
        t, ia = j // nang, j % nang
        imx, imy = t // ny + gr["min_imx"], t % ny + gr["min_imy"]
        hm["tracenum"] = j + 1
        hm["o1"], hm["n1"], hm["d1"] = 0.0, nz, gr["dz"]
        hm["imx"], hm["imy"] = imx, imy
        hm["mx"], hm["my"] = _scale(gr["dx"], imx), _scale(gr["dy"], imy)
        if nang > 1:  # ShotProfileWEM.jl:129-132
            hm["iang"], hm["iaz"] = ia // nangy, ia % nangy
            hm["ang"] = (ia // nangy) * np.float64(dangx) + oangx
            hm["az"] = (ia % nangy) * np.float64(dangy) + oangy
        fh = seis.extent((nz, nx, ny, nangx, nangy), (gr["oz"], gr["ox"], gr["oy"], oangx, oangy),
                         (gr["dz"], gr["dx"], gr["dy"], dangx, dangy), ("Depth", "mx", "my", "angx", "angy"),
                         ("m", "m", "m", "degrees", "degrees"))
        seis.write(m, image.reshape(nm * nang, nz), hm, fh)
        return image[:, 0, :] if nang == 1 else image
    mm, _, _ = seis.read(m)
    j = np.arange(nm)
    pos = (j // ny + gr["min_imx"]) * ny + j % ny + gr["min_imy"]
    if mm.shape[1] != nz or (pos.max() + 1) * nang > mm.shape[0]:
        raise ValueError(f"{m}: needs {(pos.max() + 1) * nang} traces of {nz} samples")
    if nang == 1:
        image = np.ascontiguousarray(mm[pos])
        rec = plan.model(DeviceArray((nm, nz), image), sx, sy).download()
    else:  # per shot: the image gathered from the angle bins (ShotProfileWEM.jl:240-275)
        gath = mm.reshape(-1, nang, nz)[pos]
        rec = np.zeros((sx.size, nm, nt), np.float32)
        for k in range(sx.size):
            img = angle_unbin(gath, angles[k][0], angles[k][1], **akw)
            rec[k] = plan.model(DeviceArray((nm, nz), img), sx[k:k + 1], sy[k:k + 1]).download()[0]
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
