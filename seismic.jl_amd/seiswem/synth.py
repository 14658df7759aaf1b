"""Synthetic inputs of the migration workloads (velocity, reflectivity, wavelet, geometry).

The record of every workload is produced by the library's own demigration of the reflectivity,
as the workloads prescribe; these generators only build the models.  Seeded with
numpy.random.default_rng.
"""
from __future__ import annotations

import numpy as np

from .wem import WemGeometry


def ricker(dt: float = 0.002, f0: float = 20.0) -> np.ndarray:
    """Ricker wavelet (src/Wavelets/Ricker.jl:21-29): nc = floor(1/(f0 dt)), t = dt*(-nc:nc),
    w = (1 - 2b) exp(-b), b = (pi f0 t)^2."""
    nw = 2.0 / (f0 * dt)
    nc = int(np.floor(nw / 2))
    t = dt * np.arange(-nc, nc + 1)
    b = (np.pi * f0 * t) ** 2
    return (1.0 - 2.0 * b) * np.exp(-b)


def wavelet_trace(nt: int, dt: float, f0: float) -> np.ndarray:
    """Ricker wavelet zero-padded to nt samples (the wavelet file the driver reads nt from)."""
    w = ricker(dt, f0)
    out = np.zeros(nt, np.float32)
    n = min(nt, w.size)
    out[:n] = w[:n]
    return out


def config1():
    """2-D SSF, reference CPU path: nx=256, dx=10; nz=200, dz=10; nt=1000, dt=4 ms; 2-60 Hz;
    v(z)=1500+0.5 z; reflectors at 500/1000/1500 m with amplitudes 0.1/-0.08/0.12; one shot at
    x=1280 m; 20 Hz Ricker."""
    g = WemGeometry(nt=1000, dt=0.004, nmx=256, dmx=10.0, nz=200, dz=10.0, fmin=2.0, fmax=60.0, pspi=False)
    z = np.arange(g.nz) * g.dz
    vel = np.repeat((1500.0 + 0.5 * z)[None, :], g.nmx, axis=0).astype(np.float32)
    refl = np.zeros((g.nm, g.nz), np.float32)
    for zr, a in ((500.0, 0.1), (1000.0, -0.08), (1500.0, 0.12)):
        refl[:, int(round(zr / g.dz))] = a
    wav = wavelet_trace(g.nt, g.dt, 20.0)
    return g, vel, refl, wav, np.array([1280.0], np.float32)


def blob_velocity(nx: int, nz: int, dx: float, dz: float, rng, nblobs: int = 12, v0: float = 1800.0) -> np.ndarray:
    """v0 plus Gaussian blobs A exp(-r^2/(2 sigma^2)), A~U(-800,800), sigma~U(100,300) m,
    centres uniform in the model, clipped to [1500, 4000] m/s.  Layout [nx, nz]."""
    x = np.arange(nx) * dx
    z = np.arange(nz) * dz
    X, Z = np.meshgrid(x, z, indexing="ij")
    v = np.full((nx, nz), v0)
    for _ in range(nblobs):
        a = rng.uniform(-800.0, 800.0)
        s = rng.uniform(100.0, 300.0)
        x0 = rng.uniform(0.0, x[-1])
        z0 = rng.uniform(0.0, z[-1])
        v += a * np.exp(-((X - x0) ** 2 + (Z - z0) ** 2) / (2.0 * s * s))
    return np.clip(v, 1500.0, 4000.0).astype(np.float32)


def random_reflectivity(nx: int, nz: int, rng, density: float = 0.05, amp: float = 0.1) -> np.ndarray:
    mask = rng.uniform(size=(nx, nz)) < density
    return np.where(mask, rng.uniform(-amp, amp, size=(nx, nz)), 0.0).astype(np.float32)


def config2(nshot: int = 64, seed: int = 2, nx: int = 1024, nz: int = 500, nt: int = 2000, fmax: float = 80.0):
    """2-D PSPI: nx=1024, dx=10; nz=500, dz=5; nt=2000, dt=2 ms; 4 reference velocities;
    1800 m/s + 12 Gaussian blobs; 5 % random reflectivity; 64 shots 160 m apart; 25 Hz Ricker."""
    g = WemGeometry(nt=nt, dt=0.002, nmx=nx, dmx=10.0, nz=nz, dz=5.0, fmin=0.0, fmax=fmax, nref=4, pspi=True)
    rng = np.random.default_rng(seed)
    vel = blob_velocity(g.nmx, g.nz, g.dmx, g.dz, rng)
    refl = random_reflectivity(g.nmx, g.nz, rng)
    wav = wavelet_trace(g.nt, g.dt, 25.0)
    sx = (80.0 + 160.0 * np.arange(nshot)).astype(np.float32)
    sx = sx[sx < g.nmx * g.dmx]
    return g, vel, refl, wav, sx


def config3(seed: int = 3, nx: int = 256, ny: int = 256, nz: int = 300, nt: int = 1500, fmax: float = 50.0,
            dx: float = 12.5, dz: float = 10.0):
    """3-D SSF: nx=ny=256, dx=dy=12.5; nz=300, dz=10; nt=1500, dt=4 ms; 3-50 Hz;
    v = 1600 + 0.6 z + 600 exp(-|r - c|^2 / (2 400^2)), c = (1600, 1600, 1500) m; 4 planar
    reflectors with normals at dip U(0, 30 deg), azimuth U(0, 360 deg) through random points,
    amplitudes +-U(0.05, 0.15); one shot at (1600, 1600); 15 Hz Ricker.  Smaller grids keep the
    same construction with the anomaly and the shot at the centre of the model.
    Layouts [nm, nz] with trace ix = imx * ny + imy."""
    g = WemGeometry(nt=nt, dt=0.004, nmx=nx, dmx=dx, nmy=ny, dmy=dx, nz=nz, dz=dz, fmin=3.0, fmax=fmax, pspi=False)
    rng = np.random.default_rng(seed)
    x = np.arange(nx) * dx
    y = np.arange(ny) * dx
    z = np.arange(nz) * dz
    full = nx == 256 and ny == 256 and nz == 300 and dx == 12.5 and dz == 10.0
    cx, cy, cz = (1600.0, 1600.0, 1500.0) if full else (x[nx // 2], y[ny // 2], z[nz // 2])
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    r2 = (X - cx) ** 2 + (Y - cy) ** 2 + (Z - cz) ** 2
    vel = (1600.0 + 0.6 * Z + 600.0 * np.exp(-r2 / (2.0 * 400.0 ** 2))).astype(np.float32).reshape(nx * ny, nz)
    refl = np.zeros((nx, ny, nz), np.float32)
    XX, YY = np.meshgrid(x, y, indexing="ij")
    for _ in range(4):
        th = np.deg2rad(rng.uniform(0.0, 30.0))
        ph = np.deg2rad(rng.uniform(0.0, 360.0))
        n = np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
        p0 = np.array([rng.uniform(0, x[-1]), rng.uniform(0, y[-1]), rng.uniform(0.15 * z[-1], 0.85 * z[-1])])
        amp = rng.uniform(0.05, 0.15) * rng.choice([-1.0, 1.0])
        zp = p0[2] - (n[0] * (XX - p0[0]) + n[1] * (YY - p0[1])) / n[2]
        iz = np.rint(zp / dz).astype(int)
        ok = (iz >= 0) & (iz < nz)
        ix, iy = np.nonzero(ok)
        refl[ix, iy, iz[ok]] += amp
    wav = wavelet_trace(g.nt, g.dt, 15.0)
    sx = np.array([cx if full else x[nx // 2]], np.float32)
    sy = np.array([cy if full else y[ny // 2]], np.float32)
    return g, vel, refl.reshape(nx * ny, nz), wav, sx, sy


def config4_request(rid: int, seed: int = 4, records: dict | None = None):
    """Request rid of the serving workload (seeded per request id): 80 % 2-D SSF with
    nx in {128, 256, 512, 1024}, nz in {200, 400}, nt in {1000, 2000, 4000}, dx = dz = 10 m,
    dt = 4 ms, 2-60 Hz, blob velocity (the 2-D generator); 20 % 3-D SSF with nx = ny in
    {128, 256}, nz = 200, nt = 1000, dx = dy = 12.5 m, dz = 10 m, 3-50 Hz, the 3-D generator's
    velocity.  One shot at the centre.  The records are random (the cost of a migration does not
    depend on the record values); records of one shape are shared through `records`.
    Returns (geom, vel, wav, d [1, nm, nt], sx, sy)."""
    rng = np.random.default_rng([seed, rid])
    if rng.uniform() < 0.8:
        nx = int(rng.choice([128, 256, 512, 1024]))
        nz = int(rng.choice([200, 400]))
        nt = int(rng.choice([1000, 2000, 4000]))
        g = WemGeometry(nt=nt, dt=0.004, nmx=nx, dmx=10.0, nz=nz, dz=10.0, fmin=2.0, fmax=60.0, pspi=False)
        vel = blob_velocity(nx, nz, g.dmx, g.dz, rng)
        wav = wavelet_trace(nt, g.dt, 20.0)
        sx, sy = np.array([nx // 2 * g.dmx], np.float32), None
    else:
        nx = int(rng.choice([128, 256]))
        g, vel, _, wav, sx, sy = config3(seed=int(rng.integers(1 << 30)), nx=nx, ny=nx, nz=200, nt=1000)
    key = (g.nm, g.nt)
    if records is not None and key in records:
        d = records[key]
    else:
        d = np.random.default_rng(key).standard_normal((1, g.nm, g.nt)).astype(np.float32)
        if records is not None:
            records[key] = d
    return g, vel, wav, d, sx, sy


def poisson_arrivals(n: int, rate: float = 50.0, seed: int = 4) -> np.ndarray:
    """Arrival times (s) of n requests of a Poisson process of the given rate."""
    return np.cumsum(np.random.default_rng([seed, 1 << 20]).exponential(1.0 / rate, size=n))


def config5_shots() -> np.ndarray:
    """Config-1 geometry with 32 shots at sx = 40 + 80 k."""
    return (40.0 + 80.0 * np.arange(32)).astype(np.float32)
