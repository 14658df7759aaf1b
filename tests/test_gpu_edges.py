"""GPU edge cases against the FP64 oracle: the smallest grids (one trace, nmx 2 and 3, nz 1), the largest
power-of-two transform (nkx 4096), receivers and sources below the surface with a shifted depth
origin, and zero records or a zero wavelet (an exactly zero image, for each packed PSPI body)."""
import numpy as np
import pytest

import oracle as O
import seiswem
from seiswem import synth
from conftest import rel_l2
from test_gpu_parity import oparams

pytestmark = pytest.mark.gpu
TOL = 1e-4


def geom(nx, nz, nt=128, pspi=False, **kw):
    return seiswem.WemGeometry(nt=nt, dt=0.004, nmx=nx, dmx=10.0, nz=nz, dz=10.0, fmin=2.0, fmax=40.0, nref=4,
                               pspi=pspi, **kw)


@pytest.mark.parametrize("nx,nz", [(1, 12), (2, 6), (3, 5), (8, 1)])
@pytest.mark.parametrize("pspi,mode", [(False, 0), (True, 0), (True, 1)])
def test_smallest_grids(gpu, nx, nz, pspi, mode):
    g = geom(nx, nz, pspi=pspi)
    rng = np.random.default_rng(nx * 10 + nz)
    vel = (1800.0 + 400.0 * rng.random((nx, nz))).astype(np.float32)
    wav = synth.wavelet_trace(g.nt, g.dt, 15.0)
    sx = 0.0
    d = rng.standard_normal((g.nm, g.nt)).astype(np.float32)
    m = seiswem.ShotProfileWEM(d[None], True, vel=vel, wav=wav, geom=g, sx=[sx], pspi_mode=mode)
    m_ref = O.wem(oparams(g, mode), vel, wav, sx, d=d, adj=True)
    r = rng.standard_normal((g.nm, g.nz)).astype(np.float32)
    dd = seiswem.ShotProfileWEM(r, False, vel=vel, wav=wav, geom=g, sx=[sx], pspi_mode=mode)[0]
    d_ref = O.wem(oparams(g, mode), vel, wav, sx, m=r, adj=False)
    e1, e2 = rel_l2(m, m_ref), rel_l2(dd, d_ref)
    print(f"nx={nx} nz={nz} pspi={pspi} mode={mode}: migrate {e1:.3e} model {e2:.3e}", flush=True)
    assert e1 <= TOL and e2 <= TOL


def test_largest_power_of_two_transform(gpu):
    """nmx 2048 with padx 2: nkx 4096, the largest power-of-two kernels (PSPI compat, packed)."""
    g = geom(2048, 6, nt=64, pspi=True)
    rng = np.random.default_rng(4)
    vel = synth.blob_velocity(g.nmx, g.nz, g.dmx, g.dz, rng, nblobs=3, v0=2000.0)
    wav = synth.wavelet_trace(g.nt, g.dt, 15.0)
    d = rng.standard_normal((g.nm, g.nt)).astype(np.float32)
    m = seiswem.ShotProfileWEM(d[None], True, vel=vel, wav=wav, geom=g, sx=[5000.0])
    m_ref = O.wem(oparams(g), vel, wav, 5000.0, d=d, adj=True)
    err = rel_l2(m, m_ref)
    print(f"nkx 4096: rel-L2 {err:.3e}", flush=True)
    assert err <= TOL


@pytest.mark.parametrize("pspi", [False, True])
def test_receivers_and_source_below_the_surface(gpu, pspi):
    """gz and sz below the first depth with a shifted depth origin: the receiver wavefield steps
    only from z >= gz and the image is zero above (wem.c:235-241)."""
    g = geom(64, 30, pspi=pspi, oz=50.0, gz=120.0, sz=80.0)
    rng = np.random.default_rng(9)
    vel = synth.blob_velocity(g.nmx, g.nz, g.dmx, g.dz, rng, nblobs=3, v0=2000.0)
    wav = synth.wavelet_trace(g.nt, g.dt, 15.0)
    d = rng.standard_normal((g.nm, g.nt)).astype(np.float32)
    m = seiswem.ShotProfileWEM(d[None], True, vel=vel, wav=wav, geom=g, sx=[300.0])
    m_ref = O.wem(oparams(g), vel, wav, 300.0, d=d, adj=True)
    assert rel_l2(m, m_ref) <= TOL
    z = g.oz + g.dz * np.arange(g.nz)
    assert np.all(m[:, z < g.gz] == 0)
    r = rng.standard_normal((g.nm, g.nz)).astype(np.float32)
    dd = seiswem.ShotProfileWEM(r, False, vel=vel, wav=wav, geom=g, sx=[300.0])[0]
    assert rel_l2(dd, O.wem(oparams(g), vel, wav, 300.0, m=r, adj=False)) <= TOL


# nkx 200 (mixed radix), 202 (Bluestein), 256 (power of two): each packed PSPI body
@pytest.mark.parametrize("nx", [100, 101, 128])
def test_zero_records_give_a_zero_image(gpu, nx):
    g = geom(nx, 20, pspi=True)
    vel = np.full((g.nmx, g.nz), 2000.0, np.float32)
    wav = synth.wavelet_trace(g.nt, g.dt, 15.0)
    m = seiswem.ShotProfileWEM(np.zeros((1, g.nm, g.nt), np.float32), True, vel=vel, wav=wav, geom=g, sx=[500.0])
    assert np.all(m == 0)


@pytest.mark.parametrize("nx", [100, 101, 128])
def test_zero_wavelet_gives_a_zero_image(gpu, nx):
    g = geom(nx, 20, pspi=True)
    rng = np.random.default_rng(nx)
    vel = synth.blob_velocity(g.nmx, g.nz, g.dmx, g.dz, rng, nblobs=3, v0=2000.0)
    d = rng.standard_normal((1, g.nm, g.nt)).astype(np.float32)
    m = seiswem.ShotProfileWEM(d, True, vel=vel, wav=np.zeros(g.nt, np.float32), geom=g, sx=[500.0])
    assert np.all(m == 0)
