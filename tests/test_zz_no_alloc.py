"""GPU test of swem_plan_migrate_noalloc, the allocation-free form of swem_plan_migrate: called
through the C ABI on two calls in a row (the second with fewer shots, a smaller work set and the
image accumulated), each checked against swem_plan_migrate on the same inputs to the workload's
floating-point bar (1e-4 rel-L2); a plan too large for the fixed scratch and a 3-D plan are refused
with their errors."""
import ctypes

import numpy as np
import pytest

import seiswem
from seiswem import _lib, synth
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def migrate_both(plan, d, sx, m0=None):
    """(swem_plan_migrate image, swem_plan_migrate_noalloc image); with m0 both accumulate onto it."""
    lib = seiswem.load()
    g = plan.geom
    sxa = np.ascontiguousarray(sx, np.float32)
    dd = seiswem.DeviceArray(d.shape, d)
    out = []
    for fn in (lib.swem_plan_migrate, lib.swem_plan_migrate_noalloc):
        m = seiswem.DeviceArray((g.nm, g.nz), m0 if m0 is not None else np.zeros((g.nm, g.nz), np.float32))
        _lib.check(fn(plan._p, dd.ptr, sxa.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), None, sxa.size, m.ptr,
                      int(m0 is not None), None))
        seiswem.synchronize()
        out.append(m.download())
    return out


@pytest.mark.parametrize("pspi", [True, False], ids=["pspi_compat", "ssf"])
def test_noalloc_matches_plan_migrate_on_two_calls(gpu, pspi):
    # config-2 grid, shallower: nx 1024 (nkx 2048), nz 300, nt 2000, 0-80 Hz (641 bins); one shot's
    # spectra (5.25 MB) and partial images (1.2 MB per frequency group) need several launches of
    # the 128 MiB scratch
    g, vel, refl, wav, sx = synth.config2(nshot=6)
    g = seiswem.WemGeometry(**{**g.__dict__, "nz": 300, "pspi": pspi})
    vel = np.ascontiguousarray(vel[:, :300])
    plan = seiswem.Plan(g, vel, wav)
    rng = np.random.default_rng(21)
    d = rng.standard_normal((6, g.nm, g.nt)).astype(np.float32)
    ref, got = migrate_both(plan, d, sx)
    err = rel_l2(got, ref)
    print(f"call 1 ({'PSPI' if pspi else 'SSF'}, 6 shots): rel-L2 {err:.3e}", flush=True)
    assert err <= 1e-4 and np.abs(got).max() > 0
    # second call: fewer shots, accumulated onto an existing image
    m0 = rng.standard_normal((g.nm, g.nz)).astype(np.float32)
    ref2, got2 = migrate_both(plan, d[:2].copy(), sx[:2], m0=m0)
    err2 = rel_l2(got2 - m0, ref2 - m0)
    print(f"call 2 (2 shots, accumulate): rel-L2 {err2:.3e}", flush=True)
    assert err2 <= 1e-4
    plan.close()


def test_noalloc_small_plan_second_call_smaller(gpu):
    """A small Bluestein-length plan (nx 100): everything fits one launch; then one shot."""
    g = seiswem.WemGeometry(nt=256, dt=0.004, nmx=100, dmx=10.0, nz=40, dz=10.0, fmin=2.0, fmax=40.0, nref=4,
                            pspi=True)
    rng = np.random.default_rng(3)
    vel = synth.blob_velocity(g.nmx, g.nz, g.dmx, g.dz, rng, nblobs=3, v0=2000.0)
    wav = synth.wavelet_trace(g.nt, g.dt, 15.0)
    plan = seiswem.Plan(g, vel, wav)
    d = rng.standard_normal((3, g.nm, g.nt)).astype(np.float32)
    sx = np.array([200.0, 500.0, 800.0], np.float32)
    ref, got = migrate_both(plan, d, sx)
    assert rel_l2(got, ref) <= 1e-4
    ref1, got1 = migrate_both(plan, d[1:2].copy(), sx[1:2])
    assert rel_l2(got1, ref1) <= 1e-4
    plan.close()


def test_noalloc_refuses_oversized_and_3d(gpu):
    lib = seiswem.load()
    # one shot's spectra alone: nmx 2048 x 8193 bins x 8 B = 134 MB > 128 MiB
    g = seiswem.WemGeometry(nt=4096, dt=0.0005, nmx=2048, dmx=10.0, nz=2, dz=10.0, fmin=0.0, fmax=1000.0, padt=4,
                            padx=1, pspi=False)
    s = g.sampling()
    assert g.nm * (s["ifmax"] - s["ifmin"]) * 8 > (128 << 20)
    vel = np.full((g.nm, g.nz), 2000.0, np.float32)
    plan = seiswem.Plan(g, vel, np.zeros(g.nt, np.float32))
    d = seiswem.DeviceArray((1, g.nm, g.nt))
    m = seiswem.DeviceArray((g.nm, g.nz))
    sx = np.array([100.0], np.float32)
    rc = lib.swem_plan_migrate_noalloc(plan._p, d.ptr, sx.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), None, 1,
                                       m.ptr, 0, None)
    assert rc == seiswem.SWEM_ENOMEM and "fixed scratch" in lib.swem_last_error().decode()
    plan.close()
    g3, vel3, _, wav3, sx3, sy3 = synth.config3(seed=1, nx=32, ny=32, nz=8, nt=128, fmax=20.0)
    plan3 = seiswem.Plan(g3, vel3, wav3)
    d3 = seiswem.DeviceArray((1, g3.nm, g3.nt))
    m3 = seiswem.DeviceArray((g3.nm, g3.nz))
    rc = lib.swem_plan_migrate_noalloc(plan3._p, d3.ptr, sx3.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                       sy3.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), 1, m3.ptr, 0, None)
    assert rc == seiswem.SWEM_EUNSUPPORTED
    plan3.close()
