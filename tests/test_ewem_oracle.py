"""CPU pins of the oracle's elastic WEM (ewem.c:3-741): the P/S separation in closed form, the
adjoint pair, and its low-frequency limit, where it reduces to the acoustic split-step migration."""
import numpy as np
import pytest

import oracle as O
from seiswem import synth


def test_separation_closed_forms():
    """A P plane wave polarised along its slowness (kx, kzp) separates into (1, 0), an S wave
    polarised across its slowness (-kzs, kx) into (0, 1); combination inverts separation for both
    matrix pairs; the adjoint separation is the transpose of the forward combination; below
    w = 20 rad/s the components pass through (up = uz, us2 = ux)."""
    w, vp, vs = 60.0, 2000.0, 1000.0
    for kx in (-0.02, 0.0, 0.013, 0.025):
        kzp = np.sqrt(max(w * w / vp ** 2 - kx * kx, 0.0))
        kzs = np.sqrt(max(w * w / vs ** 2 - kx * kx, 0.0))
        pp = np.array([kx, kzp]) / np.hypot(kx, kzp)
        ps = np.array([-kzs, kx]) / np.hypot(kx, kzs)
        up, us = O.esep_k(kx, w, vp, vs, True, False, complex(pp[0]), complex(pp[1]))
        assert abs(up - 1) < 1e-12 and abs(us) < 1e-12
        up, us = O.esep_k(kx, w, vp, vs, True, False, complex(ps[0]), complex(ps[1]))
        assert abs(up) < 1e-12 and abs(us - 1) < 1e-12
        for adj in (False, True):
            a1, a2 = complex(0.3, -1.2), complex(-0.7, 0.4)
            p1, p2 = O.esep_k(kx, w, vp, vs, True, adj, a1, a2)
            b1, b2 = O.esep_k(kx, w, vp, vs, False, adj, p1, p2)
            assert abs(b1 - a1) < 1e-12 and abs(b2 - a2) < 1e-12
        # transpose: <sep_adj(e_i), e_j> == <e_i, comb_fwd(e_j)>
        E = [(1.0 + 0j, 0j), (0j, 1.0 + 0j)]
        S = np.array([O.esep_k(kx, w, vp, vs, True, True, *e) for e in E]).T
        C = np.array([O.esep_k(kx, w, vp, vs, False, False, *e) for e in E]).T
        assert np.allclose(S, C.T, atol=1e-12)
    up, us = O.esep_k(0.01, 15.0, vp, vs, True, False, 2.0 + 0j, 5.0 + 0j)
    assert (up, us) == (5.0 + 0j, 2.0 + 0j)


def small(nx=48, nz=24, nt=128, fmin=0.0, fmax=30.0, seed=0):
    p = O.WemParams(nt=nt, dt=0.004, nmx=nx, dmx=10.0, nz=nz, dz=10.0, fmin=fmin, fmax=fmax, padt=2, padx=2,
                    pspi=False)
    rng = np.random.default_rng(seed)
    vp = synth.blob_velocity(nx, nz, p.dmx, p.dz, rng, nblobs=2, v0=2000.0)
    vs = (vp / 1.8).astype(np.float32)
    wav = synth.wavelet_trace(nt, p.dt, 15.0)
    return p, vp, vs, wav, rng


def test_ewem_is_an_adjoint_pair():
    """<d, L m> = <L^H d, m> over both components and both images (the separation pairs are
    transposes, the split-step pairs adjoint, injection and imaging adjoint)."""
    p, vp, vs, wav, rng = small()
    mpp, mps2 = rng.standard_normal((2, p.nm, p.nz))
    dx, dz = rng.standard_normal((2, p.nm, p.nt))
    fx, fz = O.ewem(p, vp, vs, wav, 200.0, mpp=mpp, mps2=mps2, adj=False)
    ap, a2 = O.ewem(p, vp, vs, wav, 200.0, ux=dx, uz=dz, adj=True)
    a = np.sum(dx * fx) + np.sum(dz * fz)
    b = np.sum(mpp * ap) + np.sum(mps2 * a2)
    assert abs(a - b) <= 1e-10 * abs(b)


def test_low_frequency_pp_image_is_the_acoustic_migration():
    """Below 20 rad/s there is no separation, so the P image of vertical data is the acoustic
    split-step image with the source scaled by 1 / (max|wav| / sqrt(ntfft)) and without the w^2 of
    the acoustic receiver spectra (ewem.c:116-146).  With padt = 1 and records that are one
    sinusoid of bin 1 (12.3 rad/s), only that bin is in play, so the two images differ by the one
    factor w^2 max|wav| / sqrt(ntfft).  With vs = vp the S image of horizontal data is the same."""
    nt = 128
    p = O.WemParams(nt=nt, dt=0.004, nmx=48, dmx=10.0, nz=24, dz=10.0, fmin=0.0, fmax=2.0, padt=1, padx=2,
                    pspi=False)
    rng = np.random.default_rng(1)
    vp = synth.blob_velocity(p.nmx, p.nz, p.dmx, p.dz, rng, nblobs=2, v0=2000.0)
    wav = synth.wavelet_trace(nt, p.dt, 15.0)
    t = np.arange(nt)
    a, b = rng.standard_normal((2, p.nm, 1))
    d = a * np.cos(2 * np.pi * t / nt) + b * np.sin(2 * np.pi * t / nt)
    zero = np.zeros_like(d)
    mpp, mps2 = O.ewem(p, vp, vp, wav, 200.0, ux=zero, uz=d, adj=True)
    assert np.all(mps2 == 0)
    m_ac = O.wem(p, vp, wav, 200.0, d=d, adj=True)
    w = 2 * np.pi / (nt * float(np.float32(p.dt)))  # dt is a float in the reference
    maxs = np.abs(wav).max() / np.sqrt(nt)
    assert np.linalg.norm(m_ac) > 0
    assert np.allclose(mpp * w * w * maxs, m_ac, rtol=1e-8, atol=1e-10 * np.abs(m_ac).max())
    _, s2 = O.ewem(p, vp, vp, wav, 200.0, ux=d, uz=zero, adj=True)
    assert np.allclose(s2, mpp, rtol=1e-12, atol=1e-15 * np.abs(mpp).max())


def test_ewem_float32_close_to_float64():
    p, vp, vs, wav, rng = small()
    dx, dz = rng.standard_normal((2, p.nm, p.nt)).astype(np.float32)
    a = O.ewem(p, vp, vs, wav, 200.0, ux=dx, uz=dz, adj=True)
    b = O.ewem(p, vp, vs, wav, 200.0, ux=dx, uz=dz, adj=True, f64=False)
    for x, y in zip(a, b):
        assert np.linalg.norm(x - y) / np.linalg.norm(x) < 1e-4


def small3d(nx=12, ny=10, nz=8, nt=64, fmax=30.0, seed=2):
    p = O.WemParams(nt=nt, dt=0.004, nmx=nx, dmx=10.0, nmy=ny, dmy=12.0, nz=nz, dz=10.0, fmin=0.0, fmax=fmax,
                    padt=2, padx=2, pspi=False)
    rng = np.random.default_rng(seed)
    z = np.arange(nz) * p.dz
    x = (np.arange(nx) * p.dmx)[:, None, None]
    y = (np.arange(ny) * p.dmy)[None, :, None]
    vp = (2000 + 0.5 * z[None, None, :] + 150 * np.exp(-((x - 60) ** 2 + (y - 50) ** 2) / 2500.0))
    vp = vp.reshape(nx * ny, nz).astype(np.float32)
    vs = (vp / 1.8).astype(np.float32)
    wav = synth.wavelet_trace(nt, p.dt, 15.0)
    return p, vp, vs, wav, rng


def test_ewem_3d_is_an_adjoint_pair():
    """3-D (nmy > 1): the separation acts on the 2-D transform with the matrix of kx alone
    (elastic_separate_3d), the split-step pairs are the 3-D ssop, the source sits at igx nmy + igy;
    the pair stays adjoint to 1e-10."""
    p, vp, vs, wav, rng = small3d()
    mpp, mps2 = rng.standard_normal((2, p.nm, p.nz))
    dx, dz = rng.standard_normal((2, p.nm, p.nt))
    fx, fz = O.ewem(p, vp, vs, wav, 50.0, 48.0, mpp=mpp, mps2=mps2, adj=False)
    ap, a2 = O.ewem(p, vp, vs, wav, 50.0, 48.0, ux=dx, uz=dz, adj=True)
    a = np.sum(dx * fx) + np.sum(dz * fz)
    b = np.sum(mpp * ap) + np.sum(mps2 * a2)
    assert abs(a - b) <= 1e-10 * abs(b)


def test_ewem_3d_low_frequency_pp_image_is_the_acoustic_migration():
    """The 3-D counterpart of the low-frequency limit: one sinusoid of bin 1 (below 20 rad/s, no
    separation: bin 1 is 12.3 rad/s at nt 128, dt 4 ms) as vertical data gives the 3-D acoustic
    split-step image times the factor 1 / (w^2 max|wav| / sqrt(ntfft)); the source position in y
    (sy) enters both the same way."""
    nt = 128
    p, vp, vs, wav, rng = small3d(nt=nt)
    p.fmax, p.padt = 2.0, 1
    t = np.arange(nt)
    a, b = rng.standard_normal((2, p.nm, 1))
    d = a * np.cos(2 * np.pi * t / nt) + b * np.sin(2 * np.pi * t / nt)
    zero = np.zeros_like(d)
    mpp, mps2 = O.ewem(p, vp, vp, wav, 50.0, 48.0, ux=zero, uz=d, adj=True)
    assert np.all(mps2 == 0)
    m_ac = O.wem(p, vp, wav, 50.0, 48.0, d=d, adj=True)
    w = 2 * np.pi / (nt * float(np.float32(p.dt)))
    maxs = np.abs(wav).max() / np.sqrt(nt)
    assert np.linalg.norm(m_ac) > 0
    assert np.allclose(mpp * w * w * maxs, m_ac, rtol=1e-8, atol=1e-10 * np.abs(m_ac).max())
    # a source elsewhere in y changes the image
    m_other, _ = O.ewem(p, vp, vp, wav, 50.0, 12.0, ux=zero, uz=d, adj=True)
    assert np.linalg.norm(m_other - mpp) > 1e-3 * np.linalg.norm(mpp)
