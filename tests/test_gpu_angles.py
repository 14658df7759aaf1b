"""GPU parity of the angle-domain path: swem_compute_angles against the oracle's compute_angles
(wem.c:566-992), and ShotProfileWEMAngles (ShotProfileWEM.jl:140-201, 246-275) against the
oracle's per-shot images binned on the host."""
import numpy as np
import pytest

import oracle as O
import seiswem
from seiswem import angles as A
from seiswem import synth
from conftest import rel_l2
from test_gpu_parity import oparams

pytestmark = pytest.mark.gpu
ANG_TOL = 1e-4
TOL = 1e-4
DOT_TOL = 1e-5


def case(nx, nz, nt=512, blobs=False, seed=0):
    g = seiswem.WemGeometry(nt=nt, dt=0.004, nmx=nx, dmx=10.0, nz=nz, dz=10.0, fmin=5.0, fmax=40.0, pspi=False)
    rng = np.random.default_rng(seed)
    if blobs:
        vel = synth.blob_velocity(nx, nz, g.dmx, g.dz, rng, nblobs=3, v0=2000.0)
    else:
        vel = np.full((nx, nz), 2000.0, np.float32)
    wav = synth.wavelet_trace(nt, g.dt, 20.0)
    return g, vel, wav, rng


# (nmx, nz, velocity with blobs): powers of two, mixed-radix (100, 150, 90, 200) and Bluestein
# (131, 97) lengths on both axes
SIZES = [(128, 100, False), (128, 128, True), (150, 90, True), (256, 200, False), (131, 97, True)]


@pytest.mark.parametrize("nx,nz,blobs", SIZES)
def test_compute_angles_matches_oracle(gpu, nx, nz, blobs):
    g, vel, wav, _ = case(nx, nz, blobs=blobs)
    sx = g.dmx * (nx // 2)
    ax, ay = seiswem.compute_angles(vel, wav, g, sx)
    rx, ry = O.angles(oparams(g), vel, wav, sx)
    err = rel_l2(ax, rx)
    flips = float(np.mean(np.sign(ax) != np.sign(rx)))
    print(f"angles nx={nx} nz={nz} blobs={blobs}: rel-L2 {err:.3e}, sign flips {flips:.2e}", flush=True)
    assert np.all(ay == 0) and np.all(ry == 0)
    assert err <= ANG_TOL


def test_angle_gathers_match_oracle_binning(gpu):
    """Three shots, 8 angle bins of 8 degrees from -32: the device path (migration per shot on
    the GPU, the shots' angles on the GPU, host binning) against oracle images binned with the
    same angles; then the oracle's own angles."""
    g, vel, wav, rng = case(128, 80, nt=256, blobs=True, seed=2)
    sx = np.array([320.0, 640.0, 960.0], np.float32)
    d = rng.standard_normal((3, g.nm, g.nt)).astype(np.float32)
    kw = dict(nangx=8, oangx=-32.0, dangx=8.0)
    angx = np.stack([seiswem.compute_angles(vel, wav, g, s)[0] for s in sx])
    angy = np.zeros_like(angx)
    gath = seiswem.ShotProfileWEMAngles(d, True, vel=vel, wav=wav, geom=g, sx=sx, angx=angx, angy=angy, **kw)
    ref = np.zeros_like(gath)
    for i, s in enumerate(sx):
        m_i = O.wem(oparams(g), vel, wav, float(s), d=d[i], adj=True).astype(np.float32)
        A.angle_bin(m_i, angx[i], angy[i], out=ref, **kw)
    err = rel_l2(gath, ref)
    print(f"angle gathers vs oracle images: rel-L2 {err:.3e}", flush=True)
    assert err <= TOL
    assert gath.shape == (g.nm, 8, g.nz) and np.count_nonzero(gath.sum(axis=(0, 2))) >= 4


def test_angle_gather_pair_dot_test(gpu):
    """The angle-domain migration and modelling of ShotProfileWEM.jl (same angles both ways) are
    an adjoint pair for split-step: <d, L m> = <L^H d, m>."""
    g, vel, wav, rng = case(128, 60, nt=256, seed=4)
    sx = np.array([400.0, 800.0], np.float32)
    kw = dict(nangx=6, oangx=-30.0, dangx=10.0)
    angx = np.stack([seiswem.compute_angles(vel, wav, g, s)[0] for s in sx])
    angy = np.zeros_like(angx)
    m = rng.standard_normal((g.nm, 6, g.nz)).astype(np.float32)
    d = rng.standard_normal((2, g.nm, g.nt)).astype(np.float32)
    lm = seiswem.ShotProfileWEMAngles(m, False, vel=vel, wav=wav, geom=g, sx=sx, angx=angx, angy=angy, **kw)
    lhd = seiswem.ShotProfileWEMAngles(d, True, vel=vel, wav=wav, geom=g, sx=sx, angx=angx, angy=angy, **kw)
    a = float(np.sum(d.astype(np.float64) * lm))
    b = float(np.sum(m.astype(np.float64) * lhd))
    print(f"angle pair dot: rel {abs(a - b) / abs(b):.3e}", flush=True)
    assert abs(a - b) / abs(b) <= DOT_TOL


@pytest.mark.parametrize("nx,ny,nz", [(128, 128, 16), (100, 60, 12), (64, 150, 10)])
def test_compute_angles_3d_matches_oracle(gpu, nx, ny, nz):
    """3-D (nkx = nmx, nky = nmy: the 3-D march's x- and y-passes with the receiver's conj(L)
    step, then z, x and y passes): both angle volumes against the oracle's 3-D transforms.  100, 60,
    64 and 150 are Bluestein lengths of the march and the lateral passes (arrays 256, 128, 128, 512)."""
    g = seiswem.WemGeometry(nt=64, dt=0.004, nmx=nx, dmx=10.0, nmy=ny, dmy=10.0, nz=nz, dz=10.0, fmin=5.0,
                            fmax=40.0, pspi=False)
    rng = np.random.default_rng(6)
    vel = (2000.0 + 200.0 * rng.random((g.nm, g.nz))).astype(np.float32)
    wav = synth.wavelet_trace(g.nt, g.dt, 20.0)
    sx, sy = 10.0 * (nx // 2 - 4), 10.0 * (ny // 2 + 6)  # (600, 700) at 128 x 128
    ax, ay = seiswem.compute_angles(vel, wav, g, sx, sy)
    rx, ry = O.angles(oparams(g), vel, wav, sx, sy)
    for name, a, r in (("x", ax, rx), ("y", ay, ry)):
        mag = rel_l2(np.abs(a), np.abs(r))
        agree = float(np.mean(np.sign(a) == np.sign(r)))
        print(f"angles 3-D {nx} x {ny} x {nz}, {name}: |angle| rel-L2 {mag:.3e}, signed {rel_l2(a, r):.3e}, "
              f"sign agreement {agree:.6f}", flush=True)
        # the sign is the reference's discrete left / right decision summed over frequencies
        # (wem.c:739-740, 671); where that sum is near zero (near the surface, far from the
        # source) FP32 and FP64 may decide differently: 2 of 262144 points in x at 128 x 128 x 16
        assert mag <= ANG_TOL and agree >= 0.9999


def test_compute_angles_3d_refused_off_grid(gpu):
    """nmx = 614 = 2 * 307 is neither a power of two, 2^a 3^b 5^c nor <= 512: SWEM_EUNSUPPORTED."""
    g = seiswem.WemGeometry(nt=64, dt=0.004, nmx=614, dmx=10.0, nmy=16, dmy=10.0, nz=8, dz=10.0, pspi=False)
    vel = np.full((g.nm, g.nz), 2000.0, np.float32)
    with pytest.raises(seiswem.SwemError) as e:
        seiswem.compute_angles(vel, synth.wavelet_trace(64, 0.004, 20.0), g, 80.0, 80.0)
    assert e.value.code == seiswem.SWEM_EUNSUPPORTED


def test_compute_angles_in_small_passes_is_identical(gpu):
    """A work budget that fits a few frequencies at a time (2-D) or a small chunk (3-D) splits the
    band into passes; the sums run in the same frequency order, so the angles are bit-identical."""
    g, vel, wav, _ = case(128, 100, blobs=True, seed=4)
    a1 = seiswem.compute_angles(vel, wav, g, 640.0)
    a2 = seiswem.compute_angles(vel, wav, g, 640.0, scratch_bytes=8 * g.nmx * g.nz * 12)
    assert np.array_equal(a1[0], a2[0]) and np.array_equal(a1[1], a2[1])
    g3 = seiswem.WemGeometry(nt=64, dt=0.004, nmx=128, dmx=10.0, nmy=128, dmy=10.0, nz=16, dz=10.0, fmin=5.0,
                             fmax=40.0, pspi=False)
    vel3 = np.full((g3.nm, g3.nz), 2000.0, np.float32)
    wav3 = synth.wavelet_trace(g3.nt, g3.dt, 20.0)
    b1 = seiswem.compute_angles(vel3, wav3, g3, 600.0, 700.0)
    b2 = seiswem.compute_angles(vel3, wav3, g3, 600.0, 700.0, scratch_bytes=6 * g3.nm * g3.nz * 12 * 2)
    assert np.array_equal(b1[0], b2[0]) and np.array_equal(b1[1], b2[1])
