"""GPU parity of elastic WEM (swem_ewem, ewem.c:3-741) against the FP64 oracle: migration and
modelling of both images / components, power-of-two and Bluestein transform lengths, the adjoint
pair, and the multi-shot driver."""
import numpy as np
import pytest

import oracle as O
import seiswem
from seiswem import synth
from conftest import rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-4
DOT_TOL = 1e-5


def case(nx, nz=30, nt=256, fmax=30.0, seed=0):
    g = seiswem.WemGeometry(nt=nt, dt=0.004, nmx=nx, dmx=10.0, nz=nz, dz=10.0, fmin=0.0, fmax=fmax, pspi=False,
                            padt=2, padx=2)
    p = O.WemParams(nt=nt, dt=0.004, nmx=nx, dmx=10.0, nz=nz, dz=10.0, fmin=0.0, fmax=fmax, padt=2, padx=2,
                    pspi=False)
    rng = np.random.default_rng(seed)
    vp = synth.blob_velocity(nx, nz, g.dmx, g.dz, rng, nblobs=3, v0=2200.0)
    vs = (vp / 1.8).astype(np.float32)
    wav = synth.wavelet_trace(nt, g.dt, 15.0)
    return g, p, vp, vs, wav, rng


def run(g, vp, vs, wav, sx, adj, a, b, c):
    dmy = g.dmy if g.dmy > 0 else g.dmx
    return seiswem.ewem(*(a, b, c, None, None, None) if adj else (None, None, None, a, b, c), wav, g.nt, g.ot, g.dt,
                        g.nmx, g.omx, g.dmx, g.nmy, g.omy, dmy, sx, 0.0, g.nz, g.oz, g.dz, g.gz, g.sz, vp, vs, g.fmin,
                        g.fmax, g.padt, g.padx, adj)


@pytest.mark.parametrize("nx", [64, 50, 150, 37])  # nkx 128 (power of two), 100 and 300 (mixed radix), 74 (Bluestein)
def test_ewem_migration_matches_oracle(gpu, nx):
    g, p, vp, vs, wav, rng = case(nx)
    sx = g.dmx * (nx // 3)
    ux, uz = rng.standard_normal((2, g.nm, g.nt)).astype(np.float32)
    uy = rng.standard_normal((g.nm, g.nt)).astype(np.float32)  # not used by the reference
    mpp, mps1, mps2 = run(g, vp, vs, wav, sx, True, ux, uy, uz)
    rpp, rps2 = O.ewem(p, vp, vs, wav, sx, ux=ux, uz=uz, adj=True)
    e1, e2 = rel_l2(mpp, rpp), rel_l2(mps2, rps2)
    print(f"ewem migrate nx={nx}: PP rel-L2 {e1:.3e}, PS rel-L2 {e2:.3e}", flush=True)
    assert e1 <= TOL and e2 <= TOL and np.all(mps1 == 0)


@pytest.mark.parametrize("nx", [64, 50, 37])
def test_ewem_modelling_matches_oracle(gpu, nx):
    g, p, vp, vs, wav, rng = case(nx, seed=3)
    sx = g.dmx * (nx // 2)
    mpp, mps2 = rng.standard_normal((2, g.nm, g.nz)).astype(np.float32)
    ux, uy, uz = run(g, vp, vs, wav, sx, False, mpp, np.ones_like(mpp), mps2)
    rx, rz = O.ewem(p, vp, vs, wav, sx, mpp=mpp, mps2=mps2, adj=False)
    e1, e2 = rel_l2(ux, rx), rel_l2(uz, rz)
    print(f"ewem model nx={nx}: ux rel-L2 {e1:.3e}, uz rel-L2 {e2:.3e}", flush=True)
    assert e1 <= TOL and e2 <= TOL and np.all(uy == 0)


def test_ewem_dot_product(gpu):
    g, p, vp, vs, wav, rng = case(64, seed=5)
    sx = 300.0
    mpp, mps2 = rng.standard_normal((2, g.nm, g.nz)).astype(np.float32)
    dx, dz = rng.standard_normal((2, g.nm, g.nt)).astype(np.float32)
    zero_m, zero_d = np.zeros_like(mpp), np.zeros_like(dx)
    fx, _, fz = run(g, vp, vs, wav, sx, False, mpp, zero_m, mps2)
    ap, _, a2 = run(g, vp, vs, wav, sx, True, dx, zero_d, dz)
    a = float(np.sum(dx.astype(np.float64) * fx) + np.sum(dz.astype(np.float64) * fz))
    b = float(np.sum(mpp.astype(np.float64) * ap) + np.sum(mps2.astype(np.float64) * a2))
    print(f"ewem dot: rel {abs(a - b) / abs(b):.3e}", flush=True)
    assert abs(a - b) / abs(b) <= DOT_TOL


def test_shot_profile_ewem_stacks_shots(gpu):
    """Two shots in one launch (swem_eshots): the stack against the oracle, the modelled shots
    against the same shots modelled alone."""
    g, p, vp, vs, wav, rng = case(64, seed=7)
    sx = np.array([150.0, 400.0], np.float32)
    ux, uy, uz = rng.standard_normal((3, 2, g.nm, g.nt)).astype(np.float32)
    mpp, mps1, mps2 = seiswem.ShotProfileEWEM((ux, uy, uz), True, vp=vp, vs=vs, wav=wav, geom=g, sx=sx)
    refs = [O.ewem(p, vp, vs, wav, float(s), ux=ux[i], uz=uz[i], adj=True) for i, s in enumerate(sx)]
    assert rel_l2(mpp, refs[0][0] + refs[1][0]) <= TOL and rel_l2(mps2, refs[0][1] + refs[1][1]) <= TOL
    dx, dy, dz = seiswem.ShotProfileEWEM((mpp, mps1, mps2), False, vp=vp, vs=vs, wav=wav, geom=g, sx=sx)
    assert dx.shape == (2, g.nm, g.nt) and np.all(dy == 0)
    for i, s in enumerate(sx):  # each shot of the batched launch equals the shot alone
        one = run(g, vp, vs, wav, float(s), False, mpp, mps1, mps2)
        assert np.array_equal(dx[i], one[0]) and np.array_equal(dz[i], one[2])
    alone = [run(g, vp, vs, wav, float(s), True, ux[i], uy[i], uz[i]) for i, s in enumerate(sx)]
    assert rel_l2(mpp, alone[0][0] + alone[1][0]) <= 1e-6


@pytest.mark.parametrize("adj", [True, False])
def test_ewem_own_band_matches_oracle(gpu, adj):
    """ewem.c samples its band its own way (ewem.c:43-52): ntfft = 2 truncf(padt nt / 2) and
    ifmax = trunc(fmax dt ntfft) + 1, where wem.c keeps padt nt and casts fmax to an integer first.
    padt 3, nt 255 (ntfft 764 against wem.c's 765) and fmax 30.5 give another band under each rule;
    the elastic plans follow ewem.c's, as the oracle does."""
    nx, nz, nt = 64, 24, 255
    g = seiswem.WemGeometry(nt=nt, dt=0.004, nmx=nx, dmx=10.0, nz=nz, dz=10.0, fmin=0.0, fmax=30.5, pspi=False,
                            padt=3, padx=2)
    p = O.WemParams(nt=nt, dt=0.004, nmx=nx, dmx=10.0, nz=nz, dz=10.0, fmin=0.0, fmax=30.5, padt=3, padx=2,
                    pspi=False)
    s = g.sampling()
    assert s["ntfft"] == 765 and int(2 * np.trunc(np.float32(3 * nt) / 2)) == 764
    rng = np.random.default_rng(9)
    vp = synth.blob_velocity(nx, nz, g.dmx, g.dz, rng, nblobs=3, v0=2200.0)
    vs = (vp / 1.8).astype(np.float32)
    wav = synth.wavelet_trace(nt, g.dt, 15.0)
    sx = 250.0
    if adj:
        ux, uz = rng.standard_normal((2, g.nm, g.nt)).astype(np.float32)
        mpp, _, mps2 = run(g, vp, vs, wav, sx, True, ux, np.zeros_like(ux), uz)
        rpp, rps2 = O.ewem(p, vp, vs, wav, sx, ux=ux, uz=uz, adj=True)
        e1, e2 = rel_l2(mpp, rpp), rel_l2(mps2, rps2)
    else:
        mpp, mps2 = rng.standard_normal((2, g.nm, g.nz)).astype(np.float32)
        ux, _, uz = run(g, vp, vs, wav, sx, False, mpp, np.zeros_like(mpp), mps2)
        rx, rz = O.ewem(p, vp, vs, wav, sx, mpp=mpp, mps2=mps2, adj=False)
        e1, e2 = rel_l2(ux, rx), rel_l2(uz, rz)
    print(f"ewem band of ewem.c, adj={adj}: rel-L2 {e1:.3e}, {e2:.3e}", flush=True)
    assert e1 <= TOL and e2 <= TOL


def case3d(nx, ny, nz=12, nt=128, fmax=30.0, seed=2, gz=0.0, sz=0.0):
    g = seiswem.WemGeometry(nt=nt, dt=0.004, nmx=nx, dmx=10.0, nmy=ny, dmy=12.0, nz=nz, dz=10.0, fmin=0.0, fmax=fmax,
                            pspi=False, padt=2, padx=2, gz=gz, sz=sz)
    p = O.WemParams(nt=nt, dt=0.004, nmx=nx, dmx=10.0, nmy=ny, dmy=12.0, nz=nz, dz=10.0, fmin=0.0, fmax=fmax,
                    padt=2, padx=2, pspi=False, gz=gz, sz=sz)
    rng = np.random.default_rng(seed)
    z = np.arange(nz) * g.dz
    x = (np.arange(nx) * g.dmx)[:, None, None]
    y = (np.arange(ny) * g.dmy)[None, :, None]
    xc, yc = 0.4 * nx * g.dmx, 0.6 * ny * g.dmy
    vp = 2000 + 0.5 * z[None, None, :] + 250 * np.exp(-((x - xc) ** 2 + (y - yc) ** 2) / (0.1 * (nx * g.dmx) ** 2))
    vp = vp.reshape(nx * ny, nz).astype(np.float32)
    vs = (vp / 1.8).astype(np.float32)
    wav = synth.wavelet_trace(nt, g.dt, 15.0)
    return g, p, vp, vs, wav, rng


def run3(g, vp, vs, wav, sx, sy, adj, a, b, c):
    return seiswem.ewem(*(a, b, c, None, None, None) if adj else (None, None, None, a, b, c), wav, g.nt, g.ot, g.dt,
                        g.nmx, g.omx, g.dmx, g.nmy, g.omy, g.dmy, sx, sy, g.nz, g.oz, g.dz, g.gz, g.sz, vp, vs, g.fmin,
                        g.fmax, g.padt, g.padx, adj)


# (nx, ny): nkx x nky 128 x 128 (power of two), 60 x 40 (mixed radix on both axes), 74 x 42
# (Bluestein on both axes); and receivers and source below the surface (depths above gz and sz
# not stepped)
CASES3 = [(64, 64, 0.0, 0.0), (30, 20, 0.0, 0.0), (37, 21, 0.0, 0.0), (32, 24, 25.0, 15.0)]
IDS3 = ["pow2", "mixed", "bluestein", "buried"]


@pytest.mark.parametrize("nx,ny,gz,sz", CASES3, ids=IDS3)
def test_ewem_3d_migration_matches_oracle(gpu, nx, ny, gz, sz):
    g, p, vp, vs, wav, rng = case3d(nx, ny, gz=gz, sz=sz)
    sx, sy = g.dmx * (nx // 3), g.dmy * (ny // 2)
    ux, uy, uz = rng.standard_normal((3, g.nm, g.nt)).astype(np.float32)
    mpp, mps1, mps2 = run3(g, vp, vs, wav, sx, sy, True, ux, uy, uz)
    rpp, rps2 = O.ewem(p, vp, vs, wav, sx, sy, ux=ux, uz=uz, adj=True)
    e1, e2 = rel_l2(mpp, rpp), rel_l2(mps2, rps2)
    print(f"3-D ewem migrate {nx}x{ny} gz={gz}: PP rel-L2 {e1:.3e}, PS rel-L2 {e2:.3e}", flush=True)
    assert e1 <= TOL and e2 <= TOL and np.all(mps1 == 0)
    if gz > 0:  # no image above the first receiver depth
        assert np.all(mpp.reshape(g.nm, g.nz)[:, : int(gz // g.dz) + (gz % g.dz > 0)] == 0)


@pytest.mark.parametrize("nx,ny,gz,sz", CASES3, ids=IDS3)
def test_ewem_3d_modelling_matches_oracle(gpu, nx, ny, gz, sz):
    g, p, vp, vs, wav, rng = case3d(nx, ny, seed=4, gz=gz, sz=sz)
    sx, sy = g.dmx * (nx // 2), g.dmy * (ny // 3)
    mpp, mps2 = rng.standard_normal((2, g.nm, g.nz)).astype(np.float32)
    ux, uy, uz = run3(g, vp, vs, wav, sx, sy, False, mpp, np.ones_like(mpp), mps2)
    rx, rz = O.ewem(p, vp, vs, wav, sx, sy, mpp=mpp, mps2=mps2, adj=False)
    e1, e2 = rel_l2(ux, rx), rel_l2(uz, rz)
    print(f"3-D ewem model {nx}x{ny} gz={gz}: ux rel-L2 {e1:.3e}, uz rel-L2 {e2:.3e}", flush=True)
    assert e1 <= TOL and e2 <= TOL and np.all(uy == 0)


def test_ewem_3d_dot_product(gpu):
    g, p, vp, vs, wav, rng = case3d(30, 20, seed=6)
    sx, sy = 110.0, 130.0
    mpp, mps2 = rng.standard_normal((2, g.nm, g.nz)).astype(np.float32)
    dx, dz = rng.standard_normal((2, g.nm, g.nt)).astype(np.float32)
    zero_m, zero_d = np.zeros_like(mpp), np.zeros_like(dx)
    fx, _, fz = run3(g, vp, vs, wav, sx, sy, False, mpp, zero_m, mps2)
    ap, _, a2 = run3(g, vp, vs, wav, sx, sy, True, dx, zero_d, dz)
    a = float(np.sum(dx.astype(np.float64) * fx) + np.sum(dz.astype(np.float64) * fz))
    b = float(np.sum(mpp.astype(np.float64) * ap) + np.sum(mps2.astype(np.float64) * a2))
    print(f"3-D ewem dot: rel {abs(a - b) / abs(b):.3e}", flush=True)
    assert abs(a - b) / abs(b) <= DOT_TOL


def test_shot_profile_ewem_3d_stacks_shots(gpu):
    """Two 3-D shots through swem_eshots: the stack against the oracle's two shots; the modelled
    shots equal the shots modelled alone; a source outside the grid is refused."""
    g, p, vp, vs, wav, rng = case3d(32, 24, seed=8)
    sx = np.array([90.0, 200.0], np.float32)
    sy = np.array([60.0, 150.0], np.float32)
    ux, uy, uz = rng.standard_normal((3, 2, g.nm, g.nt)).astype(np.float32)
    mpp, mps1, mps2 = seiswem.ShotProfileEWEM((ux, uy, uz), True, vp=vp, vs=vs, wav=wav, geom=g, sx=sx, sy=sy)
    refs = [O.ewem(p, vp, vs, wav, float(sx[i]), float(sy[i]), ux=ux[i], uz=uz[i], adj=True) for i in range(2)]
    assert rel_l2(mpp, refs[0][0] + refs[1][0]) <= TOL and rel_l2(mps2, refs[0][1] + refs[1][1]) <= TOL
    dx, dy, dz = seiswem.ShotProfileEWEM((mpp, mps1, mps2), False, vp=vp, vs=vs, wav=wav, geom=g, sx=sx, sy=sy)
    assert dx.shape == (2, g.nm, g.nt) and np.all(dy == 0)
    for i in range(2):
        one = run3(g, vp, vs, wav, float(sx[i]), float(sy[i]), False, mpp, mps1, mps2)
        assert np.array_equal(dx[i], one[0]) and np.array_equal(dz[i], one[2])
    with pytest.raises(seiswem.SwemError) as e:
        seiswem.ShotProfileEWEM((ux, uy, uz), True, vp=vp, vs=vs, wav=wav, geom=g, sx=sx, sy=np.array([60.0, 1e4]))
    assert e.value.code == seiswem.SWEM_EINVAL
