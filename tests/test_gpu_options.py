"""GPU tests of plan options and of concurrent calls:

* frequencies per 3-D y-pass workgroup: 1 and the 1024-thread maximum of 16 on a power-of-two grid
  against the FP64 oracle; 16 on a Bluestein-length y axis (512-thread kernels, maximum 8) and on
  3-D PSPI is refused with SWEM_EINVAL;
* SWEM_PREC_FP32 (the reference's float expressions for the kz and split-step phases,
  wem.c:336-343, 353) on 2-D SSF, 2-D PSPI and 3-D SSF against the FP64 oracle (the FP32
  oracle's distance is reported);
* several host threads calling swem_wem (wem.c:3-11) and the shotwem `main` (shotwem.c:12-140) at
  once, each result checked against the oracle: the library keeps no global planner state (the
  reference's concurrent FFTW planning is why its driver runs one process per shot,
  ShotProfileWEM.jl:104)."""
import ctypes
import threading

import numpy as np
import pytest

import oracle as O
import seiswem
from seiswem import synth
from conftest import rel_l2
from test_capi import header_array, write_seis
from test_gpu_shotwem import call_shotwem, read_seis

pytestmark = pytest.mark.gpu


def oparams(g, mode=0):
    return O.WemParams(nt=g.nt, dt=g.dt, nmx=g.nmx, dmx=g.dmx, nz=g.nz, dz=g.dz, nmy=g.nmy, dmy=g.dmy, ot=g.ot,
                       omx=g.omx, oz=g.oz, gz=g.gz, sz=g.sz, nref=g.nref, fmin=g.fmin, fmax=g.fmax, padt=g.padt,
                       padx=g.padx, pspi=g.pspi, pspi_mode=mode)


@pytest.mark.parametrize("fpb", [1, 16])
def test_3d_freqs_per_block_matches_oracle(gpu, fpb):
    g, vel, refl, wav, sx, sy = synth.config3(seed=4, nx=64, ny=64, nz=12, nt=128, fmax=25.0)
    d = np.random.default_rng(fpb).standard_normal((g.nm, g.nt)).astype(np.float32)
    m_ref = O.wem(oparams(g), vel, wav, float(sx[0]), float(sy[0]), d=d, adj=True)
    m = seiswem.ShotProfileWEM(d[None], True, vel=vel, wav=wav, geom=g, sx=sx, sy=sy, freqs_per_block=fpb)
    err = rel_l2(m, m_ref)
    print(f"3-D migration, {fpb} frequencies per y-pass workgroup: rel-L2 {err:.3e}")
    assert err <= 1e-4
    m8 = seiswem.ShotProfileWEM(d[None], True, vel=vel, wav=wav, geom=g, sx=sx, sy=sy, freqs_per_block=8)
    # the grouping changes only the order in which the y pass sums the image over frequencies
    assert rel_l2(m, m8) <= 1e-5


@pytest.mark.parametrize("nx,ny,pspi", [(50, 40, False), (64, 64, True)])
def test_3d_freqs_per_block_above_cap_is_refused(gpu, nx, ny, pspi):
    """Bluestein y lengths (ny 40 -> nky 80) and 3-D PSPI use 512-thread y-pass kernels: at most 8
    frequencies per workgroup."""
    g, vel, refl, wav, sx, sy = synth.config3(seed=4, nx=nx, ny=ny, nz=8, nt=128, fmax=20.0)
    g.pspi = pspi
    g.nref = 4
    with pytest.raises(seiswem.SwemError) as e:
        seiswem.Plan(g, vel, wav, freqs_per_block=16)
    assert e.value.code == seiswem.SWEM_EINVAL
    seiswem.Plan(g, vel, wav, freqs_per_block=8).close()


def fp32_case(kind):
    rng = np.random.default_rng(12)
    if kind == "3d":
        g, vel, refl, wav, sx, sy = synth.config3(seed=6, nx=64, ny=64, nz=16, nt=128, fmax=25.0)
    else:
        g = seiswem.WemGeometry(nt=256, dt=0.004, nmx=128, dmx=10.0, nz=40, dz=10.0, fmin=2.0, fmax=40.0, nref=4,
                                pspi=kind == "pspi")
        vel = synth.blob_velocity(g.nmx, g.nz, g.dmx, g.dz, rng, nblobs=4, v0=2000.0)
        wav = synth.wavelet_trace(g.nt, g.dt, 15.0)
        sx, sy = np.array([640.0], np.float32), np.zeros(1, np.float32)
    d = rng.standard_normal((g.nm, g.nt)).astype(np.float32)
    return g, vel, wav, sx, sy, d


@pytest.mark.parametrize("kind", ["ssf", "pspi", "3d"])
def test_naive_fp32_precision_matches_oracles(gpu, kind):
    g, vel, wav, sx, sy, d = fp32_case(kind)
    p = oparams(g)
    m64 = O.wem(p, vel, wav, float(sx[0]), float(sy[0]), d=d, adj=True)
    m32 = O.wem(p, vel, wav, float(sx[0]), float(sy[0]), d=d, adj=True, f64=False)
    m = seiswem.ShotProfileWEM(d[None], True, vel=vel, wav=wav, geom=g, sx=sx, sy=sy, precision=seiswem.PREC_FP32)
    mk = seiswem.ShotProfileWEM(d[None], True, vel=vel, wav=wav, geom=g, sx=sx, sy=sy)
    e64, e32, ek, o32 = rel_l2(m, m64), rel_l2(m, m32), rel_l2(mk, m64), rel_l2(m32, m64)
    print(f"{kind}: naive FP32 vs FP64 oracle {e64:.3e}, vs FP32 oracle {e32:.3e}; KZ64 vs FP64 oracle {ek:.3e}; "
          f"FP32 oracle vs FP64 oracle {o32:.3e}")
    # the bar is the FP64 oracle; the FP32 oracle (serial float sums, its own float FFT) is itself
    # up to 1.5e-4 from it on PSPI, so the distance to it is bounded by the two distances
    assert e64 <= 1e-4 and ek <= 1e-4
    assert e32 <= e64 + o32 + 1e-6
    assert not np.array_equal(m, mk)  # the precision option reaches the kernels


def test_concurrent_wem_calls_from_threads(gpu):
    """Four threads, each running swem_wem on its own shot (SSF and PSPI alternating) twice."""
    cases = []
    for i in range(4):
        g, vel, wav, sx, sy, d = fp32_case("pspi" if i % 2 else "ssf")
        d = d * (1.0 + 0.25 * i)
        sx = sx + 40.0 * i
        ref = O.wem(oparams(g), vel, wav, float(sx[0]), d=d, adj=True)
        cases.append((g, vel, wav, float(sx[0]), d, ref))
    out, errs = [None] * 8, []

    def run(k):
        try:
            g, vel, wav, sx, d, _ = cases[k % 4]
            out[k] = seiswem.wem(d, None, wav, g.nt, 0.0, g.dt, g.nmx, 0.0, g.dmx, 1, 0.0, g.dmx, sx, 0.0, g.nz, 0.0,
                                 g.dz, 0.0, 0.0, vel, g.nref, g.fmin, g.fmax, g.padt, g.padx, True, g.pspi)
        except Exception as e:  # reported below
            errs.append(repr(e))

    ths = [threading.Thread(target=run, args=(k,)) for k in range(8)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    assert not errs, errs
    for k in range(8):
        err = rel_l2(out[k], cases[k % 4][5])
        print(f"thread {k}: rel-L2 {err:.3e}")
        assert err <= 1e-4
        assert np.array_equal(out[k], out[(k + 4) % 8])  # same input, same image, whatever ran beside it


def test_concurrent_shotwem_main_from_threads(gpu, monkeypatch):
    """Three threads run shotwem's `main` at once on their own .seis files (absolute names, short
    enough for the 100-byte name fields of the .seis header)."""
    import pathlib
    import shutil
    import tempfile
    monkeypatch.delenv("DATAPATH", raising=False)
    tmp_path = pathlib.Path(tempfile.mkdtemp(prefix="sw", dir="/tmp"))
    lib = seiswem.load()
    nx, nz, nt, dx, dz, dt = 64, 30, 256, 10.0, 10.0, 0.004
    jobs = []
    for i in range(3):
        rng = np.random.default_rng(30 + i)
        vel = synth.blob_velocity(nx, nz, dx, dz, rng, nblobs=3, v0=2000.0)
        wav = synth.wavelet_trace(nt, dt, 15.0)
        d = rng.standard_normal((nx, nt)).astype(np.float32)
        mx = dx * np.arange(nx)
        base = tmp_path / f"job{i}"
        base.mkdir()
        f = {k: str(base / f"{k}.seis") for k in ("vel", "wav", "d", "m")}
        write_seis(lib, f["vel"], vel, header_array(nx, tracenum=np.arange(1, nx + 1), n1=nz, mx=mx,
                                                    imx=np.arange(nx)), nz, nx, o1=0.0, d1=dz, o2=0.0, d2=dx)
        write_seis(lib, f["wav"], wav[None], header_array(1, n1=nt), nt, 1, o1=0.0, d1=dt)
        write_seis(lib, f["d"], d, header_array(nx, tracenum=np.arange(1, nx + 1), n1=nt, gx=mx, sx=200.0), nt, nx,
                   o1=0.0, d1=dt, o2=0.0, d2=dx)
        pspi = i % 2 == 1
        args = ["adj=y", f"d={f['d']}", f"m={f['m']}", f"vel={f['vel']}", f"wav={f['wav']}",
                f"pspi={'y' if pspi else 'n'}", "sx=200", "fmax=40", "nref=4", "padt=2", "padx=2"]
        p = O.WemParams(nt=nt, dt=dt, nmx=nx, dmx=dx, nz=nz, dz=dz, fmin=0.0, fmax=40.0, nref=4, padt=2, padx=2,
                        pspi=pspi, pspi_mode=0)
        jobs.append((args, f["m"], O.wem(p, vel, wav, 200.0, d=d, adj=True)))
    rcs = [None] * 3
    ths = [threading.Thread(target=lambda k=k: rcs.__setitem__(k, call_shotwem(jobs[k][0]))) for k in range(3)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    assert rcs == [0, 0, 0]
    for args, mname, ref in jobs:
        m, _, _ = read_seis(mname, nz, nx)
        err = rel_l2(m, ref)
        print(f"{mname}: rel-L2 {err:.3e}")
        assert err <= 1e-4
    shutil.rmtree(tmp_path, ignore_errors=True)


def near_cutoff_case():
    """2-D split-step, v = 2000 m/s everywhere (pd = 0), nx 1024 (nkx 2048, dkx = 2 pi / 20480 m),
    nz 200, 40-44 Hz (bins 0.125 Hz apart).  The record is a sum of plane waves cos(w t - kx x) at
    the three bins whose grid wavenumbers come closest below the cutoff w / v: at 40.625 and
    43.75 Hz, w / v is exactly 416 and 448 dkx, so s = w^2 p^2 - kx^2 is zero and its float
    evaluation is rounding alone; at 42.875 Hz kx / (w / v) = 0.99991."""
    g = seiswem.WemGeometry(nt=1000, dt=0.004, nmx=1024, dmx=10.0, nz=200, dz=10.0, fmin=40.0, fmax=44.0,
                            pspi=False)
    s = g.sampling()
    vel = np.full((g.nm, g.nz), 2000.0, np.float32)
    wav = synth.wavelet_trace(g.nt, g.dt, 40.0)
    t = np.arange(g.nt) * g.dt
    x = np.arange(g.nmx) * g.dmx
    dkx = 2 * np.pi / (s["nkx"] * g.dmx)
    fb = np.arange(s["ifmin"], s["ifmax"]) / (s["ntfft"] * g.dt)
    r = 2 * np.pi * fb / 2000.0 / dkx  # w / v in units of dkx
    d = np.zeros((g.nm, g.nt))
    kxs = []
    for i in np.argsort(r - np.floor(r), kind="stable")[:3]:
        w, kx = 2 * np.pi * fb[i], np.floor(r[i]) * dkx
        kxs.append(kx / (w / 2000.0))
        d += np.cos(w * t[None, :] - kx * x[:, None])
    return g, vel, wav, d.astype(np.float32), kxs


def test_precision_near_evanescent_cutoff(gpu):
    """Per-image error against the FP64 oracle of the two GPU variants on waves near the cutoff:
    the FP64 kz tables keep the default within the 1e-4 bar, and the naive FP32 phases
    (s = (w*w)*p2 - kx*kx in float, wem.c:341) carry the cancellation into the image."""
    g, vel, wav, d, kxs = near_cutoff_case()
    p = oparams(g)
    m64 = O.wem(p, vel, wav, 5120.0, d=d, adj=True)
    mk = seiswem.ShotProfileWEM(d[None], True, vel=vel, wav=wav, geom=g, sx=[5120.0])
    mf = seiswem.ShotProfileWEM(d[None], True, vel=vel, wav=wav, geom=g, sx=[5120.0], precision=seiswem.PREC_FP32)
    m32 = O.wem(p, vel, wav, 5120.0, d=d, adj=True, f64=False)
    ek, ef, eo = rel_l2(mk, m64), rel_l2(mf, m64), rel_l2(m32, m64)
    print(f"near cutoff (kx / (w/v) = {', '.join(f'{k:.5f}' for k in kxs)}): FP64 kz {ek:.3e}, naive FP32 {ef:.3e}, "
          f"FP32 oracle {eo:.3e}")
    assert ek <= 1e-4
    assert ef > 3 * ek
