"""A seeded sweep of small random geometries through every 2-D mode and kernel choice against the
FP64 oracle: grid sizes (powers of two and Bluestein lengths, padx 1 and 2), depth origins, source
and receiver depths, bands, source positions and velocities drawn from a fixed generator."""
import numpy as np
import pytest

import oracle as O
import seiswem
from seiswem import synth
from conftest import rel_l2
from test_gpu_parity import oparams

pytestmark = pytest.mark.gpu
TOL = 1e-4


def draw(i):
    rng = np.random.default_rng(1000 + i)
    nx = int(rng.choice([16, 37, 50, 64, 90, 128, 200, 256]))
    padx = int(rng.choice([1, 2]))
    nz = int(rng.integers(4, 25))
    nt = int(rng.choice([64, 100, 128, 200]))
    dt = float(rng.choice([0.002, 0.004]))
    pspi = bool(rng.integers(0, 2))
    mode = int(rng.integers(0, 2)) if pspi else 0
    variant = int(rng.choice([seiswem.VARIANT_AUTO, seiswem.VARIANT_PAIR, seiswem.VARIANT_SUBGROUP,
                              seiswem.VARIANT_WAVE, seiswem.VARIANT_PACKED]))
    dz = float(rng.choice([5.0, 10.0]))
    oz = float(rng.choice([0.0, 20.0]))
    sz = float(rng.choice([0.0, oz + 2 * dz]))
    gz = float(rng.choice([0.0, oz + dz]))
    fmax = float(rng.choice([25.0, 40.0, 60.0]))
    g = seiswem.WemGeometry(nt=nt, dt=dt, nmx=nx, dmx=10.0, nz=nz, dz=dz, oz=oz, sz=sz, gz=gz, fmin=0.0,
                            fmax=fmax, nref=int(rng.integers(2, 6)), pspi=pspi, padx=padx, padt=2)
    vel = (1500.0 + 1500.0 * rng.random((nx, nz))).astype(np.float32)
    if rng.integers(0, 2):
        vel = np.repeat(vel[:1], nx, axis=0)  # laterally constant: the homogeneous-depth branches
    wav = synth.wavelet_trace(nt, dt, float(rng.choice([10.0, 20.0])))
    sx = float(10.0 * rng.integers(0, nx))
    return g, vel, wav, sx, mode, variant, rng


@pytest.mark.parametrize("i", range(24))
def test_random_geometry(gpu, i):
    g, vel, wav, sx, mode, variant, rng = draw(i)
    d = rng.standard_normal((g.nm, g.nt)).astype(np.float32)
    m = seiswem.ShotProfileWEM(d[None], True, vel=vel, wav=wav, geom=g, sx=[sx], pspi_mode=mode, variant=variant)
    m_ref = O.wem(oparams(g, mode), vel, wav, sx, d=d, adj=True)
    r = rng.standard_normal((g.nm, g.nz)).astype(np.float32)
    dd = seiswem.ShotProfileWEM(r, False, vel=vel, wav=wav, geom=g, sx=[sx], pspi_mode=mode, variant=variant)[0]
    d_ref = O.wem(oparams(g, mode), vel, wav, sx, m=r, adj=False)
    e1 = rel_l2(m, m_ref) if np.linalg.norm(m_ref) > 0 else float(np.abs(m).max())
    e2 = rel_l2(dd, d_ref) if np.linalg.norm(d_ref) > 0 else float(np.abs(dd).max())
    print(f"case {i}: nx={g.nmx} padx={g.padx} nz={g.nz} pspi={g.pspi} mode={mode} variant={variant}: "
          f"migrate {e1:.2e} model {e2:.2e}", flush=True)
    assert e1 <= TOL and e2 <= TOL


def draw3(i):
    rng = np.random.default_rng(2000 + i)
    nx = int(rng.choice([8, 20, 33, 50, 64, 100, 128]))
    ny = int(rng.choice([4, 9, 16, 37, 64]))
    padx = int(rng.choice([1, 2]))
    nz = int(rng.integers(3, 11))
    nt = int(rng.choice([64, 128]))
    pspi = bool(rng.integers(0, 2))
    dz = float(rng.choice([5.0, 10.0]))
    oz = float(rng.choice([0.0, 20.0]))
    sz = float(rng.choice([0.0, oz + 2 * dz]))
    gz = float(rng.choice([0.0, oz + dz]))
    g = seiswem.WemGeometry(nt=nt, dt=0.004, nmx=nx, dmx=10.0, nmy=ny, dmy=10.0, nz=nz, dz=dz, oz=oz, sz=sz, gz=gz,
                            fmin=2.0, fmax=float(rng.choice([20.0, 35.0])), nref=int(rng.integers(2, 5)), pspi=pspi,
                            padx=padx, padt=2)
    vel = (1500.0 + 1500.0 * rng.random((nx * ny, nz))).astype(np.float32)
    if rng.integers(0, 2):
        vel = np.repeat(vel[:1], nx * ny, axis=0)
    wav = synth.wavelet_trace(nt, g.dt, 15.0)
    sx, sy = float(10.0 * rng.integers(0, nx)), float(10.0 * rng.integers(0, ny))
    return g, vel, wav, sx, sy, rng


@pytest.mark.parametrize("i", range(12))
def test_random_geometry_3d(gpu, i):
    """3-D (nmy > 1; PSPI in the reference's compat mode) on random power-of-two and Bluestein
    lengths of both axes."""
    g, vel, wav, sx, sy, rng = draw3(i)
    s = g.sampling()
    d = rng.standard_normal((g.nm, g.nt)).astype(np.float32)
    m = seiswem.ShotProfileWEM(d[None], True, vel=vel, wav=wav, geom=g, sx=[sx], sy=[sy])
    m_ref = O.wem(oparams(g, seiswem.PSPI_COMPAT), vel, wav, sx, sy, d=d, adj=True)
    r = rng.standard_normal((g.nm, g.nz)).astype(np.float32)
    dd = seiswem.ShotProfileWEM(r, False, vel=vel, wav=wav, geom=g, sx=[sx], sy=[sy])[0]
    d_ref = O.wem(oparams(g, seiswem.PSPI_COMPAT), vel, wav, sx, sy, m=r, adj=False)
    e1 = rel_l2(m, m_ref) if np.linalg.norm(m_ref) > 0 else float(np.abs(m).max())
    e2 = rel_l2(dd, d_ref) if np.linalg.norm(d_ref) > 0 else float(np.abs(dd).max())
    print(f"3-D case {i}: {g.nmx}x{g.nmy} nkx {s['nkx']} nky {s['nky']} nz={g.nz} pspi={g.pspi}: "
          f"migrate {e1:.2e} model {e2:.2e}", flush=True)
    assert e1 <= TOL and e2 <= TOL
