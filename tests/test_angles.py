"""CPU tests of the angle-domain path: the oracle's compute_angles (wem.c:566-992) pinned by
geometry, and the bilinear angle binning of ShotProfileWEM.jl:162-188 / 246-275."""
import numpy as np
import pytest

import oracle as O
from seiswem import angles as A
from seiswem import synth


def homogeneous(nx=128, nz=100, nmy=1, nt=512):
    """Constant 2000 m/s.  nmx > 100 turns the reference's edge taper on (wem.c:792-795), which
    with nkx = nmx (no padding in compute_angles) keeps waves from wrapping around in x."""
    p = O.WemParams(nt=nt, dt=0.004, nmx=nx, dmx=10.0, nz=nz, dz=10.0, nmy=nmy, dmy=10.0, fmin=5, fmax=40,
                    pspi=False)
    vel = np.full((p.nm, nz), 2000.0, np.float32)
    wav = synth.wavelet_trace(p.nt, p.dt, 20.0)
    return p, vel, wav


def test_oracle_angles_match_point_source_geometry():
    """In a constant velocity the source wavefield of a surface point source travels along rays
    from the source, so |du/dx| / |du/dz| = |x - sx| / z: the angle is atan(|x - sx| / z) in
    degrees, negative on one side of the source and positive on the other (a wrong derivative
    axis, a missing atan or degree conversion, or a sign mix-up fails this)."""
    p, vel, wav = homogeneous()
    sx = 640.0
    ax, ay = O.angles(p, vel, wav, sx)
    assert np.all(ay == 0.0)  # nky = 1: ky = 0 everywhere
    x = np.arange(p.nmx) * p.dmx
    z = (np.arange(p.nz) + 1) * p.dz  # image row iz holds the wavefield after step iz
    geo = np.degrees(np.arctan(np.abs(x[:, None] - sx) / z[None, :]))
    # away from the 30-trace edge tapers and from the source's near field
    ix = np.arange(p.nmx)[:, None]
    sel = (z[None, :] >= 150.0) & (geo <= 60.0) & (np.abs(x[:, None] - sx) >= 100.0) & (ix >= 30) & (ix < 98)
    err = np.abs(np.abs(ax) - geo)[sel]
    assert err.mean() < 1.0 and np.percentile(err, 99) < 5.0 and err.max() < 10.0
    side = np.sign(x[:, None] - sx) * np.ones_like(geo)
    assert np.all(np.sign(ax[sel]) == side[sel])
    # odd about the source trace igx = 64 (up to the taper, symmetric about 63.5, and the Nyquist
    # bin, which the reference counts as positive kx) below the near-surface rows
    d = np.abs(ax[64 - 20:64][::-1] + ax[65:65 + 20])[:, 10:]
    assert d.mean() < 0.5 and d.max() < 3.0


def test_oracle_angles_float32_close_to_float64():
    p, vel, wav = homogeneous()
    a64, _ = O.angles(p, vel, wav, 640.0)
    a32, _ = O.angles(p, vel, wav, 640.0, f64=False)
    assert np.linalg.norm(a32 - a64) / np.linalg.norm(a64) < 5e-4


def test_oracle_angles_3d_axes():
    """3-D: with the source at the centre of a square grid the y angles are the x angles with
    x and y exchanged (a transposed axis or a ky / kx mix-up fails this) and angx is odd in
    x - sx.  (The grid is too small for the ray-geometry check of the 2-D test: without padding
    the field wraps around within the record; the x path itself is the one that test pins.)"""
    n, nz = 32, 24
    p, vel, wav = homogeneous(n, nz, nmy=n, nt=128)
    sx = sy = 160.0
    ax, ay = O.angles(p, vel, wav, sx, sy)
    A, B = ax.reshape(n, n, nz), ay.reshape(n, n, nz)
    assert np.abs(A - B.transpose(1, 0, 2)).max() <= 1e-9 * np.abs(A).max()
    iy = 16  # y = sy
    # odd about sx (on the source's y line, below the near-surface rows) up to the Nyquist bin,
    # which the reference counts as positive kx
    lft, rgt = A[16 - 6:16, iy, 10:][::-1], A[17:17 + 6, iy, 10:]
    assert np.abs(np.abs(lft) - np.abs(rgt)).max() < 1.0
    assert (np.sign(lft) == -np.sign(rgt)).mean() > 0.95  # sign decisions in the wrapped field may tie



def test_angle_bin_weights_closed_form():
    nm, nz = 3, 2
    m = np.arange(1, nm * nz + 1, dtype=np.float32).reshape(nm, nz)
    ax = np.array([[0.0, 2.5], [4.0, 9.9], [-1.0, 7.0]])
    ay = np.zeros_like(ax)
    out = A.angle_bin(m, ax, ay, nangx=5, oangx=0.0, dangx=2.0)
    # 0.0 -> bin 0 weight 1; 2.5 -> bins 1, 2 with 0.75, 0.25; 4.0 -> bin 2; 9.9 -> bin 4 (the
    # +x corner dropped past the last bin); -1.0 -> outside; 7.0 -> bins 3, 4 with 0.5, 0.5
    want = np.zeros((nm, 5, nz), np.float32)
    want[0, 0, 0] = 1
    want[0, 1, 1], want[0, 2, 1] = 0.75 * 2, 0.25 * 2
    want[1, 2, 0] = 3
    want[1, 4, 1] = (10.0 - 9.9) / 2.0 * 4
    want[2, 3, 1], want[2, 4, 1] = 0.5 * 6, 0.5 * 6
    assert np.allclose(out, want, atol=1e-6)


@pytest.mark.parametrize("nangy", [1, 3])
def test_angle_bin_unbin_are_transposes(nangy):
    rng = np.random.default_rng(0)
    nm, nz, nangx = 7, 9, 6
    ax = rng.uniform(-10, 70, (nm, nz))
    ay = rng.uniform(-5, 40, (nm, nz)) if nangy > 1 else np.zeros((nm, nz))
    kw = dict(nangx=nangx, oangx=0.0, dangx=10.0, nangy=nangy, oangy=0.0, dangy=12.0)
    m = rng.standard_normal((nm, nz)).astype(np.float32)
    g = rng.standard_normal((nm, nangx * nangy, nz)).astype(np.float32)
    a = float(np.sum(A.angle_bin(m, ax, ay, **kw).astype(np.float64) * g))
    b = float(np.sum(m.astype(np.float64) * A.angle_unbin(g, ax, ay, **kw)))
    assert abs(a - b) <= 1e-5 * max(abs(a), 1.0)


def test_single_bin_is_plain_stack():
    """nangx = nangy = 1 with zero angles puts the whole shot image in the one bin."""
    rng = np.random.default_rng(1)
    m = rng.standard_normal((5, 4)).astype(np.float32)
    z = np.zeros((5, 4))
    out = A.angle_bin(m, z, z, nangx=1, oangx=0.0, dangx=1.0)
    assert np.array_equal(out[:, 0, :], m)
