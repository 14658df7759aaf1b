"""Pins of the CPU oracle (test infrastructure) against closed forms, textbook identities and
the reference's own arithmetic.  The reference ships no golden vectors for this path and cannot
be built here (it needs FFTW3), so each oracle function is pinned by a check a plausible
mistake would fail:

  FFT               brute-force DFT (n <= 64, all radices, both signs), numpy.fft (large n,
                    prime n), inverse(forward(x)) = n x
  f_op r2c / c2r    numpy.fft.rfft, impulse -> all-ones spectrum, c2r(r2c(x)) = n x, DC and
                    Nyquist imaginary parts ignored
  taper             closed-form values at i = 0, l-1, l, n-l, n-l+1, n-1; identity for n <= 100
  sampling          the configuration table values with the reference's float casts
  source index      (int)(sx-omx)/dmx truncation order
  PSPI tables       a hand-worked ladder; v(z) gives iref 0 and no cut
  ssop step         plane wave e^{i kx x} -> e^{i kz dz} (propagating) / e^{-0.2 sqrt|s| |dz|}
                    (evanescent), times the taper when nmx > 100
  wem adjoint/fwd   dot test <= 1e-12 (SSF, PSPI linear); PSPI compat is not adjoint (as the
                    reference); linearity; shot superposition; PSPI == SSF for v(z) without
                    taper; 1-D kinematics t0 = t_wavelet + 2 sum dz/v; config-1 reflector
                    depths and polarities
"""
import numpy as np
import pytest

import oracle as O
from conftest import rel_l2

rng0 = np.random.default_rng(1234)


def dft(x, sign):
    n = x.size
    k = np.arange(n)
    return np.exp(sign * 2j * np.pi * np.outer(k, k) / n) @ x


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 15, 16, 25, 27, 30, 32, 49, 60, 64])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft_brute_force(n, sign):
    x = rng0.standard_normal(n) + 1j * rng0.standard_normal(n)
    ref = dft(x, sign)
    assert np.max(np.abs(O.fft(x, sign) - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))
    assert np.max(np.abs(O.fft(x, sign, f64=False) - ref)) <= 2e-6 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("n", [128, 512, 1000, 1009, 2000, 2048, 3000, 4000])
def test_fft_numpy_and_roundtrip(n):
    x = rng0.standard_normal(n) + 1j * rng0.standard_normal(n)
    y = O.fft(x, -1)
    assert rel_l2(y, np.fft.fft(x)) <= 1e-14
    assert rel_l2(O.fft(y, +1) / n, x) <= 1e-14
    assert rel_l2(O.fft(x, -1, f64=False), np.fft.fft(x)) <= 1e-6


def test_fop_r2c_c2r():
    nt, nw = 300, 401  # f_op length (nw-1)*2 = 800
    n = 2 * (nw - 1)
    x = rng0.standard_normal(nt)
    X = O.fop(x, nw, nt, adj=True)
    assert rel_l2(X, np.fft.rfft(np.concatenate([x, np.zeros(n - nt)]))) <= 1e-14
    imp = np.zeros(nt)
    imp[0] = 1.0
    assert np.allclose(O.fop(imp, nw, nt, adj=True), 1.0, atol=1e-14)
    back = O.fop(X, nw, nt, adj=False)
    assert rel_l2(back, n * x) <= 1e-13
    X2 = X.copy()
    X2[0] += 5j
    X2[-1] += 3j
    assert np.allclose(O.fop(X2, nw, nt, adj=False), back, atol=1e-9)


def test_taper_closed_form():
    n, l = 256, 30
    t = O.taper(n, l)
    assert t[0] == pytest.approx(np.exp(-(0.015 * 30) ** 2), rel=1e-15)
    assert t[l - 1] == pytest.approx(np.exp(-(0.015) ** 2), rel=1e-15)
    assert np.all(t[l:n - l + 1] == 1.0)
    assert t[n - l + 1] == pytest.approx(np.exp(-(0.015) ** 2), rel=1e-15)
    assert t[n - 1] == pytest.approx(np.exp(-(0.015 * (l - 1)) ** 2), rel=1e-15)
    assert np.allclose(O.taper(n, l, f64=False), t, rtol=2e-7)
    assert np.all(O.taper(64, 0) == 1.0)


@pytest.mark.parametrize("cfg,expect", [
    (dict(nt=1000, dt=0.004, nmx=256, dmx=10, nz=200, dz=10, fmin=2, fmax=60), (2000, 1001, 512, 1, 16, 481, 30)),
    (dict(nt=2000, dt=0.002, nmx=1024, dmx=10, nz=500, dz=5, fmin=0, fmax=80), (4000, 2001, 2048, 1, 0, 641, 30)),
    (dict(nt=1500, dt=0.004, nmx=256, nmy=256, dmx=12.5, dmy=12.5, nz=300, dz=10, fmin=3, fmax=50),
     (3000, 1501, 512, 512, 36, 601, 30)),
    # the (int) cast binds to fmax/fmin only: 60.9 Hz behaves as 60 Hz, 2.9 Hz as 2 Hz
    (dict(nt=1000, dt=0.004, nmx=256, dmx=10, nz=200, dz=10, fmin=2.9, fmax=60.9), (2000, 1001, 512, 1, 16, 481, 30)),
    # padt=1 (the C default): no time padding; fmax beyond Nyquist clamps to nw
    (dict(nt=500, dt=0.004, nmx=50, dmx=10, nz=20, dz=10, fmin=0, fmax=200, padt=1, padx=1), (500, 251, 50, 1, 0, 251, 0)),
])
def test_sampling_config_values(cfg, expect):
    p = O.WemParams(**cfg)
    s = O.sampling(p)
    assert (s.ntfft, s.nw, s.nkx, s.nky, s.ifmin, s.ifmax, s.lmx) == expect


def test_source_index_truncation_order():
    p = O.WemParams(nt=10, dt=0.004, nmx=256, dmx=10, nz=10, dz=10)
    assert O.source_index(p, 1280.0) == (0, 128, 0)
    assert O.source_index(p, 1289.9)[1] == 128       # (int)1289.9 = 1289, /10 = 128.9 -> 128
    p.omx = 3.0
    assert O.source_index(p, 1282.7)[1] == 127       # (int)1279.7 = 1279 -> 127.9 -> 127
    assert O.source_index(p, 5000.0)[0] != 0          # outside the grid is rejected


def test_pspi_tables_hand_worked():
    p = O.WemParams(nt=10, dt=0.004, nmx=4, dmx=10, nz=2, dz=10, nref=4, pspi=True)
    vel = np.array([[1500, 2000], [2000, 2000], [2400, 2000], [1800, 2000]], np.float32)
    vref, i1, i2, cut = O.pspi_tables(p, vel)
    assert np.array_equal(vref[:, 0], np.array([1500, 1800, 2100, 2400], np.float32))
    assert list(i1[:, 0]) == [0, 1, 3, 1] and list(i2[:, 0]) == [1, 2, 3, 2]
    assert list(cut[:, 0]) == [1, 1, 0, 1]
    # laterally invariant layer: vmax == vmin, every point uses vref 0 without the real-part cut
    assert np.all(i1[:, 1] == 0) and np.all(i2[:, 1] == 0) and np.all(cut[:, 1] == 0)


def test_pspi_vmin_over_first_row_only_is_clamped():
    """3-D: vmin is searched over ix < nmx only (wem.c:49); a slower point elsewhere would give
    iref < 0 in the reference; the oracle clamps it to 0."""
    p = O.WemParams(nt=10, dt=0.004, nmx=2, nmy=2, dmx=10, dmy=10, nz=1, dz=10, nref=3, pspi=True)
    vel = np.array([[2000], [2400], [2200], [1000]], np.float32)
    vref, i1, i2, cut = O.pspi_tables(p, vel)
    assert vref[0, 0] == 2000 and vref[-1, 0] == 2400
    assert i1[3, 0] == 0


def plane_wave_case(nmx, v, f_bin, kx_bin, padx=1):
    p = O.WemParams(nt=64, dt=0.004, nmx=nmx, dmx=10.0, nz=3, dz=10.0, fmin=0, fmax=125, padx=padx, pspi=False)
    vel = np.full((nmx, 3), v, np.float32)
    x = np.arange(nmx)
    u = np.exp(2j * np.pi * kx_bin * x / nmx)
    s = O.sampling(p)
    # dt and dmx reach the oracle as floats (wem.c:3-11 takes float)
    w = 2 * np.pi * f_bin / (s.ntfft * float(np.float32(p.dt)))
    kx = 2 * np.pi * kx_bin / (nmx * float(np.float32(p.dmx)))
    return p, vel, u, w, kx


@pytest.mark.parametrize("kx_bin", [0, 3, 10])
def test_ssf_plane_wave_propagating(kx_bin):
    p, vel, u, w, kx = plane_wave_case(64, 2000.0, 20, kx_bin)
    kz = np.sqrt((w / 2000.0) ** 2 - kx ** 2)
    for dzs in (10.0, -10.0):
        out = O.step(p, vel, u, 20, 1, dzs, adj=True)
        assert np.max(np.abs(out - u * np.exp(1j * kz * dzs))) <= 1e-12


def test_ssf_plane_wave_evanescent_and_taper():
    p, vel, u, w, kx = plane_wave_case(128, 2000.0, 3, 30)
    s = (w / 2000.0) ** 2 - kx ** 2
    assert s < 0
    out = O.step(p, vel, u, 3, 0, 10.0, adj=True)
    expect = u * np.exp(-0.2 * np.sqrt(-s) * 10.0) * O.taper(128, 30)
    assert np.max(np.abs(out - expect)) <= 1e-12


def lateral_case(nmx=64, pspi=False, mode=0, nref=4, f_bin=20, iz=1):
    """padx = 1, nmx <= 100 (no taper, nkx = nmx): a laterally constant wavefield is the kx = 0
    mode alone, so one step multiplies it by closed-form phases even over a laterally varying
    velocity.  v(x) runs 2000 -> 2600 m/s, chosen off the reference-velocity ladder's steps."""
    p = O.WemParams(nt=64, dt=0.004, nmx=nmx, dmx=10.0, nz=3, dz=10.0, fmin=0, fmax=125, padx=1, pspi=pspi,
                    pspi_mode=mode, nref=nref)
    x = np.arange(nmx)
    v = (2000.0 + 600.0 * x / (nmx - 1)).astype(np.float32)
    vel = np.repeat(v[:, None], 3, 1)
    s = O.sampling(p)
    w = 2 * np.pi * f_bin / (s.ntfft * float(np.float32(p.dt)))
    u = np.full(nmx, 0.7 - 0.4j)
    return p, vel, v.astype(np.float64), u, w


@pytest.mark.parametrize("dzs", [10.0, -10.0])
def test_ssf_split_step_phase_closed_form(dzs):
    """ssop over v(x) (wem.c:313-359).  Adjoint branch: FFT, kz phase at kx = 0 (exp(i w po dz)),
    IFFT, then the split-step factor exp(i w pd dz): a constant u leaves as u exp(i w dz / v(x)).
    Forward branch (adj = 0, wem.c:321-326): the split-step factor comes first, so the input
    c exp(-i w pd dz) is constant after it and leaves as c exp(i w po dz).  A flipped sign of the
    split-step phase, a dropped pd, or po taken from another depth fails both."""
    p, vel, v, u, w = lateral_case()
    out = O.step(p, vel, u, 20, 1, dzs, adj=True)
    assert np.max(np.abs(out - u * np.exp(1j * w * dzs / v))) <= 1e-12
    po = 1.0 / np.mean(v)
    pd = 1.0 / v - po
    uf = u * np.exp(-1j * w * pd * dzs)
    outf = O.step(p, vel, uf, 20, 1, dzs, adj=False)
    assert np.max(np.abs(outf - u * np.exp(1j * w * po * dzs))) <= 1e-12
    # the naive FP32 restatement follows within float rounding of the phase
    out32 = O.step(p, vel, u, 20, 1, dzs, adj=True, f64=False)
    assert np.max(np.abs(out32 - u * np.exp(1j * w * dzs / v))) <= 2e-5


def vref_ladder(v, nref):
    """vref[i] = vmin + i (vmax - vmin) / (nref - 1) and iref = trunc((nref - 1)(v - vmin) /
    (vmax - vmin)), both clamped to nref - 1 above nref - 2 (wem.c:39-70)."""
    vmin, vmax = float(v.min()), float(v.max())
    vref = vmin + np.arange(nref) * (vmax - vmin) / (nref - 1)
    iref = np.floor((nref - 1) * (v - vmin) / (vmax - vmin)).astype(int)
    i1 = np.where(iref > nref - 2, nref - 1, iref)
    i2 = np.where(iref > nref - 2, nref - 1, iref + 1)
    return vref, i1, i2


@pytest.mark.parametrize("dzs", [5.0, -5.0])
def test_pspi_compat_selection_closed_form(dzs):
    """pspiop compat (wem.c:428-459, linear_interp wem.c:473-478) on the kx = 0 mode over v(x):
    each vref propagates the constant u to u exp(i w dz / vref), and point x keeps the result of
    vref[iref1(x)] -- only its real part where vref2 - vref1 > 10 (every point below vmax), the
    full complex value where iref1 == iref2 (v = vmax).  Taking iref2, Im instead of Re, or
    skipping the cut fails."""
    nref = 4
    p, vel, v, u, w = lateral_case(pspi=True, mode=0, nref=nref)
    vref, i1, i2 = vref_ladder(v, nref)
    cut = (vref[i2] - vref[i1]) > 10
    assert cut.sum() == v.size - 1 and not cut[-1] and np.all(i2[cut] == i1[cut] + 1)
    y1 = u * np.exp(1j * w * dzs / vref[i1])
    expect = np.where(cut, y1.real, y1)
    for adj in (True, False):  # pspiop ignores adj (wem.c:380-471)
        out = O.step(p, vel, u, 20, 1, dzs, adj=adj)
        assert np.max(np.abs(out - expect)) <= 1e-12
    out32 = O.step(p, vel, u, 20, 1, dzs, adj=True, f64=False)
    assert np.max(np.abs(out32 - expect)) <= 2e-5


def test_pspi_linear_weights_closed_form():
    """pspi_mode 1, select after propagation (the adjoint branch): cut points take
    (1 - w2) y1 + w2 y2 with w2 = (v - vref1) / (vref2 - vref1), the interpolation of
    wem.c:476.  The forward branch (weight per x, propagate, sum) is pinned as the transpose of
    the adjoint one: L is even in kx, so F^-1 L F = F L F^-1 and every factor is symmetric."""
    nref = 4
    p, vel, v, u, w = lateral_case(pspi=True, mode=1, nref=nref)
    vref, i1, i2 = vref_ladder(v, nref)
    cut = (vref[i2] - vref[i1]) > 10
    w2 = np.where(cut, (v - vref[i1]) / np.where(cut, vref[i2] - vref[i1], 1.0), 0.0)
    assert np.all((w2 >= 0) & (w2 < 1)) and w2.max() > 0.5
    y1 = u * np.exp(1j * w * 10.0 / vref[i1])
    y2 = u * np.exp(1j * w * 10.0 / vref[i2])
    expect = np.where(cut, (1 - w2) * y1 + w2 * y2, y1)
    out = O.step(p, vel, u, 20, 1, 10.0, adj=True)
    assert np.max(np.abs(out - expect)) <= 1e-12
    # forward branch = the transpose of the adjoint one: <P_fwd a, b> = <a, P_adj^T b> for the
    # complex bilinear form, since every factor (weights, phases, unitary FFTs) is symmetric
    rng = np.random.default_rng(11)
    a = rng.standard_normal(v.size) + 1j * rng.standard_normal(v.size)
    b = rng.standard_normal(v.size) + 1j * rng.standard_normal(v.size)
    lhs = np.sum(O.step(p, vel, a, 20, 1, 10.0, adj=False) * b)
    rhs = np.sum(a * O.step(p, vel, b, 20, 1, 10.0, adj=True))
    assert abs(lhs - rhs) <= 1e-12 * abs(rhs)


def pspi_lateral_images(amp, nrefs, nx=48, nz=20, nt=128):
    base = O.WemParams(nt=nt, dt=0.004, nmx=nx, dmx=10.0, nz=nz, dz=10.0, fmin=2, fmax=30, pspi=False)
    z = np.arange(nz) * base.dz
    x = np.arange(nx)[:, None] * base.dmx
    vel = (2000 + 0.5 * z[None, :] + amp * np.sin(2 * np.pi * x / 470.0)).astype(np.float32)
    wav = np.zeros(nt, np.float32)
    wav[:25] = np.hanning(25)
    d = np.random.default_rng(8).standard_normal((nx, nt))
    m_ssf = O.wem(base, vel, wav, 240.0, d=d, adj=True)
    ms = {n: O.wem(O.WemParams(**{**base.__dict__, "pspi": True, "pspi_mode": 1, "nref": n}), vel, wav, 240.0,
                   d=d, adj=True) for n in nrefs}
    return m_ssf, ms


def test_pspi_linear_ladder_convergence():
    """Linear PSPI (wem.c:476's interpolation) converges at second order in the ladder step while
    the steps stay above the 10 m/s cut threshold (contrast 480 m/s, nref 3 -> 33): halving the
    step divides the distance to the nref = 33 image by more than 2.5.  Weights swapped between
    vref1 and vref2, or nearest-vref selection, converge at first order (factor 2) or not at all."""
    _, ms = pspi_lateral_images(240.0, (3, 5, 9, 17, 33))
    errs = [rel_l2(ms[n], ms[33]) for n in (3, 5, 9)]
    assert errs[0] / errs[1] > 2.5 and errs[1] / errs[2] > 2.5


def test_pspi_and_ssf_agree_to_first_order_in_contrast():
    """PSPI and split-step share the phase-shift operator of v(z) and differ at first order in the
    lateral contrast: quartering the contrast quarters their distance (ratio 3.5..4.5; the
    cut threshold stays inactive: nref 4 over a 120+ m/s range).  A split-step correction of the
    wrong sign or size, or a wrong PSPI reference velocity, leaves a contrast-independent part."""
    d = []
    for amp in (240.0, 60.0):
        m_ssf, ms = pspi_lateral_images(amp, (4,))
        d.append(rel_l2(ms[4], m_ssf))
    assert 3.5 <= d[0] / d[1] <= 4.5


def small(pspi, mode=0, lateral=True, seed=0, nx=48, nz=24, nt=128):
    p = O.WemParams(nt=nt, dt=0.004, nmx=nx, dmx=10.0, nz=nz, dz=10.0, fmin=2, fmax=50, nref=4, pspi=pspi,
                    pspi_mode=mode)
    rng = np.random.default_rng(seed)
    z = np.arange(nz) * p.dz
    vel = np.repeat((1600 + 0.8 * z)[None, :], nx, 0)
    if lateral:
        x = np.arange(nx)[:, None] * p.dmx
        vel = vel + 400 * np.exp(-((x - 200) ** 2 + (z[None, :] - 120) ** 2) / (2 * 80.0 ** 2))
    wav = np.zeros(nt, np.float32)
    wav[:25] = np.hanning(25)
    return p, vel.astype(np.float32), wav, rng


@pytest.mark.parametrize("pspi,mode", [(False, 0), (True, 1)])
def test_dot_product_fp64(pspi, mode):
    p, vel, wav, rng = small(pspi, mode)
    d = rng.standard_normal((p.nm, p.nt))
    m = rng.standard_normal((p.nm, p.nz))
    a = np.sum(d * O.wem(p, vel, wav, 200.0, m=m, adj=False))
    b = np.sum(m * O.wem(p, vel, wav, 200.0, d=d, adj=True))
    assert abs(a - b) / abs(b) <= 1e-12


def test_pspi_compat_not_adjoint():
    p, vel, wav, rng = small(True, 0)
    d = rng.standard_normal((p.nm, p.nt))
    m = rng.standard_normal((p.nm, p.nz))
    a = np.sum(d * O.wem(p, vel, wav, 200.0, m=m, adj=False))
    b = np.sum(m * O.wem(p, vel, wav, 200.0, d=d, adj=True))
    assert abs(a - b) / abs(b) > 1e-3


def test_linearity_and_shot_superposition():
    p, vel, wav, rng = small(True, 0, seed=3)
    d1, d2 = rng.standard_normal((2, p.nm, p.nt))
    m1 = O.wem(p, vel, wav, 150.0, d=d1)
    m2 = O.wem(p, vel, wav, 150.0, d=d2)
    assert rel_l2(O.wem(p, vel, wav, 150.0, d=2.5 * d1 - d2), 2.5 * m1 - m2) <= 1e-12
    ms = O.shots(p, vel, wav, [150.0, 300.0], d=np.stack([d1, d2]))
    assert rel_l2(ms, m1 + O.wem(p, vel, wav, 300.0, d=d2)) <= 1e-13


def test_pspi_equals_ssf_for_vz_without_taper():
    """Laterally invariant v(z), nmx <= 100: PSPI (vref 0 = v, no cut) and SSF (pd = 0) are the
    same phase-shift migration, reached through different code paths."""
    p, vel, wav, rng = small(False, 0, lateral=False)
    d = rng.standard_normal((p.nm, p.nt))
    m_ssf = O.wem(p, vel, wav, 200.0, d=d, adj=True)
    p.pspi = True
    m_pspi = O.wem(p, vel, wav, 200.0, d=d, adj=True)
    assert rel_l2(m_pspi, m_ssf) <= 1e-12


def test_one_d_kinematics():
    """nmx = 1: kx = 0 and each step is exp(i w p dz).  A record impulse at t0 images at the
    depth where 2 sum(dz/v) = t0 - t_w, t_w the wavelet's centre."""
    v, dz, nz, nt, dt = 2000.0, 10.0, 60, 512, 0.002
    p = O.WemParams(nt=nt, dt=dt, nmx=1, dmx=10.0, nz=nz, dz=dz, fmin=0, fmax=120, pspi=False)
    vel = np.full((1, nz), v, np.float32)
    f0 = 30.0
    nc = int(np.floor(1.0 / (f0 * dt)))
    t = dt * np.arange(-nc, nc + 1)
    b = (np.pi * f0 * t) ** 2
    wav = np.zeros(nt, np.float32)
    wav[: t.size] = (1 - 2 * b) * np.exp(-b)
    tw = nc * dt
    for iz0 in (15, 30, 44):
        t0 = tw + 2 * (iz0 + 1) * dz / v
        d = np.zeros((1, nt))
        d[0, int(round(t0 / dt))] = 1.0
        m = O.wem(p, vel, wav, 0.0, d=d, adj=True)[0]
        assert int(np.argmax(m)) == iz0


def test_config1_kinematics_fp64():
    nx, nz = 256, 200
    p = O.WemParams(nt=1000, dt=0.004, nmx=nx, dmx=10, nz=nz, dz=10, fmin=2, fmax=60, pspi=False)
    z = np.arange(nz) * 10.0
    vel = np.repeat((1500 + 0.5 * z)[None, :], nx, 0).astype(np.float32)
    m = np.zeros((nx, nz))
    m[:, 50], m[:, 100], m[:, 150] = 0.1, -0.08, 0.12
    f0, dt = 20.0, 0.004
    nc = int(np.floor(1.0 / (f0 * dt)))
    t = dt * np.arange(-nc, nc + 1)
    b = (np.pi * f0 * t) ** 2
    wav = np.zeros(1000, np.float32)
    wav[: t.size] = (1 - 2 * b) * np.exp(-b)
    d = O.wem(p, vel, wav, 1280.0, m=m, adj=False)
    img = O.wem(p, vel, wav, 1280.0, d=d, adj=True)
    prof = img[128]
    for iz, sgn in ((50, 1), (100, -1), (150, 1)):
        win = prof[iz - 3: iz + 4]
        assert np.argmax(np.abs(win)) == 3 and np.sign(win[3]) == sgn
    img32 = O.wem(p, vel, wav, 1280.0, d=d.astype(np.float32), adj=True, f64=False)
    assert rel_l2(img32, img) <= 1e-4


# ----------------------------------------------------------------------------- 3-D (nmy > 1)
def plane_wave_3d(nmx, nmy, dmx, dmy, v, f_bin, kx_bin, ky_bin):
    p = O.WemParams(nt=64, dt=0.004, nmx=nmx, dmx=dmx, nmy=nmy, dmy=dmy, nz=3, dz=10.0, fmin=0, fmax=125, padx=1,
                    pspi=False)
    vel = np.full((nmx * nmy, 3), v, np.float32)
    X, Y = np.meshgrid(np.arange(nmx), np.arange(nmy), indexing="ij")
    u = np.exp(2j * np.pi * (kx_bin * X / nmx + ky_bin * Y / nmy)).reshape(-1)
    s = O.sampling(p)
    w = 2 * np.pi * f_bin / (s.ntfft * float(np.float32(p.dt)))
    kx = 2 * np.pi * kx_bin / (nmx * float(np.float32(dmx)))
    ky = 2 * np.pi * ky_bin / (nmy * float(np.float32(dmy)))
    return p, vel, u, w, kx, ky


@pytest.mark.parametrize("kx_bin,ky_bin", [(0, 0), (3, 0), (0, 5), (4, -3)])
def test_ssf_plane_wave_3d(kx_bin, ky_bin):
    """An on-grid 3-D plane wave exp(i (kx x + ky y)) is multiplied by exp(i kz dz) with
    kz = sqrt(w^2/v^2 - kx^2 - ky^2): a swapped dkx/dky or a dropped ky term fails it
    (dmx != dmy, nmx != nmy)."""
    p, vel, u, w, kx, ky = plane_wave_3d(32, 16, 10.0, 20.0, 2000.0, 20, kx_bin, ky_bin)
    kz = np.sqrt((w / 2000.0) ** 2 - kx ** 2 - ky ** 2)
    for dzs in (10.0, -10.0):
        out = O.step(p, vel, u, 20, 1, dzs, adj=True)
        assert np.max(np.abs(out - u * np.exp(1j * kz * dzs))) <= 1e-12


def small3d_oracle(nx=24, ny=16, nz=10, nt=64, dmx=10.0, dmy=15.0, seed=0):
    p = O.WemParams(nt=nt, dt=0.004, nmx=nx, dmx=dmx, nmy=ny, dmy=dmy, nz=nz, dz=10.0, fmin=3, fmax=40, pspi=False)
    rng = np.random.default_rng(seed)
    X, Y, Z = np.meshgrid(np.arange(nx) * dmx, np.arange(ny) * dmy, np.arange(nz) * 10.0, indexing="ij")
    vel = (1700 + 0.5 * Z + 300 * np.exp(-((X - 100) ** 2 + (Y - 120) ** 2 + (Z - 50) ** 2) / (2 * 60.0 ** 2)))
    wav = np.zeros(nt, np.float32)
    wav[:15] = np.hanning(15)
    return p, vel.reshape(nx * ny, nz).astype(np.float32), wav, rng


def test_dot_product_fp64_3d():
    p, vel, wav, rng = small3d_oracle()
    d = rng.standard_normal((p.nm, p.nt))
    m = rng.standard_normal((p.nm, p.nz))
    a = np.sum(d * O.wem(p, vel, wav, 100.0, 120.0, m=m, adj=False))
    b = np.sum(m * O.wem(p, vel, wav, 100.0, 120.0, d=d, adj=True))
    assert abs(a - b) / abs(b) <= 1e-12


def test_xy_transposition_3d():
    """Swapping the roles of x and y (sizes, spacings, velocity, data and source position)
    transposes the image: a transposed operand in the 2-D transform or the trace index fails it."""
    p, vel, wav, rng = small3d_oracle()
    nx, ny, nz = p.nmx, p.nmy, p.nz
    d = rng.standard_normal((p.nm, p.nt))
    m = O.wem(p, vel, wav, 100.0, 120.0, d=d, adj=True)
    q = O.WemParams(nt=p.nt, dt=p.dt, nmx=ny, dmx=p.dmy, nmy=nx, dmy=p.dmx, nz=nz, dz=p.dz, fmin=p.fmin, fmax=p.fmax,
                    pspi=False)
    tr = lambda a: a.reshape(nx, ny, -1).transpose(1, 0, 2).reshape(nx * ny, -1)  # noqa: E731
    mt = O.wem(q, tr(vel), wav, 120.0, 100.0, d=tr(d), adj=True)
    assert np.max(np.abs(mt - tr(m))) <= 1e-10 * np.max(np.abs(m))
