"""GPU parity of the 3-D split-step path (nmy > 1) through the C ABI against the FP64 oracle.

Tolerances as the 2-D tests: relative L2 <= 1e-4 against the FP64 oracle on the same seeded
inputs, dot-product tests <= 1e-5.  The full config-3 shape (256 x 256 x 300, 565 frequencies)
is checked through the dot test and linearity, whose results do not depend on an oracle run.
"""
import numpy as np
import pytest

import oracle as O
import seiswem
from seiswem import synth
from conftest import rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-4
DOT_TOL = 1e-5


def oparams(g: seiswem.WemGeometry) -> O.WemParams:
    return O.WemParams(nt=g.nt, dt=g.dt, nmx=g.nmx, dmx=g.dmx, nz=g.nz, dz=g.dz, nmy=g.nmy, dmy=g.dmy, ot=g.ot,
                       omx=g.omx, oz=g.oz, gz=g.gz, sz=g.sz, nref=g.nref, fmin=g.fmin, fmax=g.fmax, padt=g.padt,
                       padx=g.padx, pspi=g.pspi, pspi_mode=0)


def small3d(nx=64, ny=64, nz=24, nt=128, fmax=25.0, seed=3):
    return synth.config3(seed=seed, nx=nx, ny=ny, nz=nz, nt=nt, fmax=fmax)


@pytest.mark.parametrize("nx,ny,nz", [(64, 64, 24), (128, 64, 12), (64, 128, 12)])
def test_migrate3d_matches_oracle(gpu, nx, ny, nz):
    """nx = 128 > 100 switches the taper on along x (lmx = 30), ny = 128 along y."""
    g, vel, refl, wav, sx, sy = small3d(nx, ny, nz)
    rng = np.random.default_rng(nx + ny)
    d = rng.standard_normal((g.nm, g.nt)).astype(np.float32)
    m_ref = O.wem(oparams(g), vel, wav, float(sx[0]), float(sy[0]), d=d, adj=True)
    m_gpu = seiswem.ShotProfileWEM(d[None], True, vel=vel, wav=wav, geom=g, sx=sx, sy=sy)
    err = rel_l2(m_gpu, m_ref)
    print(f"3-D migrate {nx}x{ny}x{nz}: rel-L2 {err:.3e}", flush=True)
    assert err <= TOL


def test_migrate3d_source_below_surface(gpu):
    """sz = 25 m: the source is not stepped at the first three depths while the receiver is
    (wem.c:235-241); the image there uses the source as injected."""
    g, vel, refl, wav, sx, sy = small3d(64, 64, 12)
    g.sz = 25.0
    rng = np.random.default_rng(5)
    d = rng.standard_normal((g.nm, g.nt)).astype(np.float32)
    m_ref = O.wem(oparams(g), vel, wav, float(sx[0]), float(sy[0]), d=d, adj=True)
    m_gpu = seiswem.ShotProfileWEM(d[None], True, vel=vel, wav=wav, geom=g, sx=sx, sy=sy)
    err = rel_l2(m_gpu, m_ref)
    print(f"3-D migrate, source below the surface: rel-L2 {err:.3e}", flush=True)
    assert err <= TOL


@pytest.mark.parametrize("nx,ny,nz", [(64, 64, 24), (128, 64, 12)])
def test_model3d_matches_oracle(gpu, nx, ny, nz):
    g, vel, refl, wav, sx, sy = small3d(nx, ny, nz)
    d_ref = O.wem(oparams(g), vel, wav, float(sx[0]), float(sy[0]), m=refl, adj=False)
    d_gpu = seiswem.ShotProfileWEM(refl, False, vel=vel, wav=wav, geom=g, sx=sx, sy=sy)[0]
    err = rel_l2(d_gpu, d_ref)
    print(f"3-D model {nx}x{ny}x{nz}: rel-L2 {err:.3e}", flush=True)
    assert err <= TOL


def dot_test(g, vel, wav, sx, sy, seed, **kw):
    plan = seiswem.Plan(g, vel, wav, **kw)
    rng = np.random.default_rng(seed)
    m = rng.standard_normal((g.nm, g.nz)).astype(np.float32)
    d = rng.standard_normal((1, g.nm, g.nt)).astype(np.float32)
    lm = plan.model(seiswem.DeviceArray(m.shape, m), sx, sy).download()
    lhd = plan.migrate(seiswem.DeviceArray(d.shape, d), sx, sy).download()
    a = float(np.sum(d.astype(np.float64) * lm))
    b = float(np.sum(m.astype(np.float64) * lhd))
    return a, b, abs(a - b) / abs(b)


def test_dot_product_3d_small(gpu):
    g, vel, refl, wav, sx, sy = small3d(128, 64, 16)
    a, b, e = dot_test(g, vel, wav, sx, sy, 5)
    print(f"3-D dot small: {a:.9e} {b:.9e} rel {e:.3e}", flush=True)
    assert e <= DOT_TOL


def test_frequency_chunks_3d(gpu):
    """A work budget below one chunk of the whole band splits it; images and records agree."""
    g, vel, refl, wav, sx, sy = small3d(64, 64, 16)
    rng = np.random.default_rng(1)
    d = rng.standard_normal((2, g.nm, g.nt)).astype(np.float32)
    sxs = np.array([sx[0], sx[0] - 200.0], np.float32)
    sys_ = np.array([sy[0], sy[0] + 100.0], np.float32)
    whole = seiswem.Plan(g, vel, wav)
    small = seiswem.Plan(g, vel, wav, scratch_bytes=2 << 20)
    dd = seiswem.DeviceArray(d.shape, d)
    m1 = whole.migrate(dd, sxs, sys_).download()
    m2 = small.migrate(dd, sxs, sys_).download()
    e2 = small.migrate_each(dd, sxs, sys_).download()
    assert rel_l2(m2, m1) <= 1e-6
    assert rel_l2(m2, e2.astype(np.float64).sum(0)) <= 1e-6
    mm = seiswem.DeviceArray(refl.shape, refl)
    r1 = whole.model(mm, sxs, sys_).download()
    r2 = small.model(mm, sxs, sys_).download()
    assert rel_l2(r2, r1) <= 1e-6


def test_config3_full_size_dot_and_linearity(gpu):
    """Config 3 at full size (nkx = nky = 512, nz = 300, 565 frequencies): the adjoint
    dot-product test and linearity of the migration."""
    g, vel, refl, wav, sx, sy = synth.config3()
    a, b, e = dot_test(g, vel, wav, sx, sy, 7)
    print(f"config3 dot: {a:.9e} {b:.9e} rel {e:.3e}", flush=True)
    assert e <= DOT_TOL
    plan = seiswem.Plan(g, vel, wav)
    rng = np.random.default_rng(8)
    d1 = rng.standard_normal((1, g.nm, g.nt)).astype(np.float32)
    d2 = rng.standard_normal((1, g.nm, g.nt)).astype(np.float32)
    m1 = plan.migrate(seiswem.DeviceArray(d1.shape, d1), sx, sy).download()
    m2 = plan.migrate(seiswem.DeviceArray(d2.shape, d2), sx, sy).download()
    m12 = plan.migrate(seiswem.DeviceArray(d1.shape, 0.5 * d1 - d2), sx, sy).download()
    err = rel_l2(m12, 0.5 * m1.astype(np.float64) - m2)
    print(f"config3 linearity rel-L2 {err:.3e}", flush=True)
    assert err <= 1e-5


def pspi3d(nx, ny, nz, lateral=True, sz=0.0, gz=0.0):
    g, vel, refl, wav, sx, sy = small3d(nx, ny, nz)
    g.pspi, g.nref, g.sz, g.gz = True, 4, sz, gz
    if not lateral:  # v(z): vmax <= vmin + 10 at every depth, pspiop keeps the complex values
        vel = np.repeat(vel.reshape(nx * ny, nz)[:1], nx * ny, 0)
    return g, vel, refl, wav, sx, sy


@pytest.mark.parametrize("nx,ny,nz,lateral,sz", [(64, 64, 16, True, 0.0), (128, 64, 10, True, 0.0),
                                                 (64, 64, 12, False, 0.0), (64, 64, 12, True, 25.0)])
def test_pspi3d_migrate_matches_oracle(gpu, nx, ny, nz, lateral, sz):
    """3-D PSPI (compat, pspiop wem.c:380-471 with nky > 1): the per-vref x-pass and the selecting
    y-pass against the FP64 oracle; a v(z) model keeps complex values (no cut), a source below the
    surface leaves the first depths unstepped for the source."""
    g, vel, refl, wav, sx, sy = pspi3d(nx, ny, nz, lateral, sz=sz)
    rng = np.random.default_rng(nx + ny + nz)
    d = rng.standard_normal((g.nm, g.nt)).astype(np.float32)
    m_ref = O.wem(oparams(g), vel, wav, float(sx[0]), float(sy[0]), d=d, adj=True)
    m_gpu = seiswem.ShotProfileWEM(d[None], True, vel=vel, wav=wav, geom=g, sx=sx, sy=sy)
    err = rel_l2(m_gpu, m_ref)
    print(f"3-D PSPI migrate {nx}x{ny}x{nz} lateral={lateral} sz={sz}: rel-L2 {err:.3e}", flush=True)
    assert err <= TOL


@pytest.mark.parametrize("nx,ny,nz,gz", [(64, 64, 16, 0.0), (64, 64, 12, 25.0)])
def test_pspi3d_model_matches_oracle(gpu, nx, ny, nz, gz):
    g, vel, refl, wav, sx, sy = pspi3d(nx, ny, nz, gz=gz)
    d_ref = O.wem(oparams(g), vel, wav, float(sx[0]), float(sy[0]), m=refl, adj=False)
    d_gpu = seiswem.ShotProfileWEM(refl, False, vel=vel, wav=wav, geom=g, sx=sx, sy=sy)[0]
    err = rel_l2(d_gpu, d_ref)
    print(f"3-D PSPI model {nx}x{ny}x{nz} gz={gz}: rel-L2 {err:.3e}", flush=True)
    assert err <= TOL


def oparams_mode(g, mode):
    p = oparams(g)
    p.pspi_mode = mode
    return p


@pytest.mark.parametrize("nx,ny,nz,sz,gz", [(64, 64, 12, 0.0, 0.0), (128, 64, 8, 25.0, 0.0), (50, 40, 8, 0.0, 25.0)])
def test_pspi3d_linear_matches_oracle(gpu, nx, ny, nz, sz, gz):
    """3-D PSPI in the linear mode (the interpolation of wem.c:476): migration selects after the
    per-vref propagation with the weights (1 - w2, w2); modelling marches the source the same way
    and the receiver with the weights applied before propagation (the exact adjoint), the y-pass
    writing one weighted slice per vref and the x-pass summing their propagated spectra.  A source
    or receiver level below the surface and Bluestein lengths (50 x 40: nkx 100, nky 80) included."""
    g, vel, refl, wav, sx, sy = pspi3d(nx, ny, nz, sz=sz, gz=gz)
    rng = np.random.default_rng(nx * ny + nz)
    d = rng.standard_normal((g.nm, g.nt)).astype(np.float32)
    po = oparams_mode(g, 1)
    m_ref = O.wem(po, vel, wav, float(sx[0]), float(sy[0]), d=d, adj=True)
    m_gpu = seiswem.ShotProfileWEM(d[None], True, vel=vel, wav=wav, geom=g, sx=sx, sy=sy,
                                   pspi_mode=seiswem.PSPI_LINEAR)
    d_ref = O.wem(po, vel, wav, float(sx[0]), float(sy[0]), m=refl, adj=False)
    d_gpu = seiswem.ShotProfileWEM(refl, False, vel=vel, wav=wav, geom=g, sx=sx, sy=sy,
                                   pspi_mode=seiswem.PSPI_LINEAR)[0]
    em, ed = rel_l2(m_gpu, m_ref), rel_l2(d_gpu, d_ref)
    print(f"3-D PSPI linear {nx}x{ny}x{nz} sz={sz} gz={gz}: migrate {em:.3e}, model {ed:.3e}", flush=True)
    assert em <= TOL and ed <= TOL


def test_pspi3d_linear_dot_product(gpu):
    """The linear-mode pair is an exact adjoint: <d, L m> = <L^H d, m> within 1e-5 (compat is not,
    as the reference)."""
    g, vel, refl, wav, sx, sy = pspi3d(64, 64, 12)
    a, b, e = dot_test(g, vel, wav, sx, sy, 11, pspi_mode=seiswem.PSPI_LINEAR)
    print(f"3-D PSPI linear dot: {a:.9e} {b:.9e} rel {e:.3e}", flush=True)
    assert e <= DOT_TOL
    _, _, ec = dot_test(g, vel, wav, sx, sy, 11)
    assert ec > 1e-3


# 3-D grids whose padded lengths (padx 2) are not powers of two 128..1024: an axis of length
# 2^a 3^b 5^c takes the mixed-radix transform on the next power of two (up to 1024), another
# length Bluestein's identity on a power-of-two array length M >= 2n - 1 (up to n = 512)
BLUE3 = [
    pytest.param(37, 41, 12, id="blue74x82"),        # nkx 74 (M 256), nky 82 (M 256)
    pytest.param(64, 37, 12, id="pow2x_blue74"),     # nkx 128 (a power of two), nky 74 (M 256)
    pytest.param(151, 32, 8, id="blue302x_mixed64"), # nkx 302 (M 1024, taper on x), nky 64 (mixed, M 128)
    pytest.param(30, 131, 8, id="mixed60x_blue262"), # nkx 60 (mixed, M 128), nky 262 (M 1024, taper on y)
    pytest.param(50, 40, 12, id="mixed100x80"),      # nkx 100, nky 80 (mixed, M 128)
    pytest.param(150, 32, 8, id="mixed300x64"),      # nkx 300 (M 512), nky 64
    pytest.param(300, 20, 6, id="mixed600x40"),      # nkx 600 (M 1024): beyond Bluestein's 512
    pytest.param(30, 125, 8, id="mixed60x250"),      # nkx 60, nky 250 (M 256)
]


@pytest.mark.parametrize("nx,ny,nz", BLUE3)
def test_migrate3d_bluestein(gpu, nx, ny, nz):
    g, vel, refl, wav, sx, sy = small3d(nx, ny, nz)
    rng = np.random.default_rng(nx * ny)
    d = rng.standard_normal((g.nm, g.nt)).astype(np.float32)
    m_ref = O.wem(oparams(g), vel, wav, float(sx[0]), float(sy[0]), d=d, adj=True)
    m_gpu = seiswem.ShotProfileWEM(d[None], True, vel=vel, wav=wav, geom=g, sx=sx, sy=sy)
    err = rel_l2(m_gpu, m_ref)
    print(f"3-D non-power-of-two migrate {nx}x{ny}x{nz}: rel-L2 {err:.3e}", flush=True)
    assert err <= TOL


@pytest.mark.parametrize("nx,ny,nz", BLUE3)
def test_model3d_bluestein(gpu, nx, ny, nz):
    g, vel, refl, wav, sx, sy = small3d(nx, ny, nz)
    d_ref = O.wem(oparams(g), vel, wav, float(sx[0]), float(sy[0]), m=refl, adj=False)
    d_gpu = seiswem.ShotProfileWEM(refl, False, vel=vel, wav=wav, geom=g, sx=sx, sy=sy)[0]
    err = rel_l2(d_gpu, d_ref)
    print(f"3-D non-power-of-two model {nx}x{ny}x{nz}: rel-L2 {err:.3e}", flush=True)
    assert err <= TOL


@pytest.mark.parametrize("nx,ny", [(37, 41), (50, 40)], ids=["blue", "mixed"])
def test_dot_product_3d_bluestein(gpu, nx, ny):
    g, vel, refl, wav, sx, sy = small3d(nx, ny, 12)
    a, b, e = dot_test(g, vel, wav, sx, sy, 9)
    print(f"3-D {nx}x{ny} dot: {a:.9e} {b:.9e} rel {e:.3e}", flush=True)
    assert e <= DOT_TOL


@pytest.mark.parametrize("nx,ny,nz", [(37, 41, 10), (50, 40, 10), (150, 32, 8)])
def test_pspi3d_bluestein(gpu, nx, ny, nz):
    """3-D PSPI (compat) on Bluestein (74 x 82) and mixed-radix lengths, migration and modelling
    against the FP64 oracle."""
    g, vel, refl, wav, sx, sy = pspi3d(nx, ny, nz)
    rng = np.random.default_rng(nx + ny)
    d = rng.standard_normal((g.nm, g.nt)).astype(np.float32)
    m_ref = O.wem(oparams(g), vel, wav, float(sx[0]), float(sy[0]), d=d, adj=True)
    m_gpu = seiswem.ShotProfileWEM(d[None], True, vel=vel, wav=wav, geom=g, sx=sx, sy=sy)
    d_ref = O.wem(oparams(g), vel, wav, float(sx[0]), float(sy[0]), m=refl, adj=False)
    d_gpu = seiswem.ShotProfileWEM(refl, False, vel=vel, wav=wav, geom=g, sx=sx, sy=sy)[0]
    e1, e2 = rel_l2(m_gpu, m_ref), rel_l2(d_gpu, d_ref)
    print(f"3-D PSPI {nx}x{ny}x{nz}: migrate {e1:.3e} model {e2:.3e}", flush=True)
    assert e1 <= TOL and e2 <= TOL


def test_3d_lengths_beyond_bluestein_are_refused(gpu):
    """nky = 2 * 307 = 614 is neither a power of two, 2^a 3^b 5^c nor <= 512: SWEM_EUNSUPPORTED."""
    g, vel, refl, wav, sx, sy = small3d(64, 307, 4)
    with pytest.raises(seiswem.SwemError) as e:
        seiswem.Plan(g, vel, wav)
    assert e.value.code == seiswem.SWEM_EUNSUPPORTED


@pytest.mark.parametrize("ny,pspi", [(40, False), (128, False), (40, True)])
def test_single_inline_3d(gpu, ny, pspi):
    """nmx = 1 with nmy > 1: nkx = 1 (wem.c:80), the x-pass transforms one point (the mixed-radix
    transform with no pass); migration and modelling against the FP64 oracle."""
    g, vel, refl, wav, sx, sy = pspi3d(1, ny, 8) if pspi else small3d(1, ny, 8)
    rng = np.random.default_rng(ny)
    d = rng.standard_normal((g.nm, g.nt)).astype(np.float32)
    m_ref = O.wem(oparams(g), vel, wav, float(sx[0]), float(sy[0]), d=d, adj=True)
    m_gpu = seiswem.ShotProfileWEM(d[None], True, vel=vel, wav=wav, geom=g, sx=sx, sy=sy)
    d_ref = O.wem(oparams(g), vel, wav, float(sx[0]), float(sy[0]), m=refl, adj=False)
    d_gpu = seiswem.ShotProfileWEM(refl, False, vel=vel, wav=wav, geom=g, sx=sx, sy=sy)[0]
    e1, e2 = rel_l2(m_gpu, m_ref), rel_l2(d_gpu, d_ref)
    print(f"3-D single inline 1 x {ny} pspi={pspi}: migrate {e1:.3e} model {e2:.3e}", flush=True)
    assert e1 <= TOL and e2 <= TOL
