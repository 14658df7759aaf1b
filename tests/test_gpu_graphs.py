"""GPU tests of the captured-graph 3-D migration (swem_plan_set_graphs): each shot's launches
(every frequency chunk and depth) replayed as one graph give the same image, bit for bit, as the
same kernels launched one by one; replays are interleaved with eager work and synchronisations
and every replay is checked; the graph is recaptured when the launch shape changes (a smaller
work budget splits the band into chunks); PSPI and per-shot images too; and the graph path
matches the FP64 oracle."""
import numpy as np
import pytest

import oracle as O
import seiswem
from seiswem import synth
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def case(pspi=False, nx=64, ny=64, nz=16, seed=3):
    g, vel, refl, wav, sx, sy = synth.config3(seed=seed, nx=nx, ny=ny, nz=nz, nt=128, fmax=25.0)
    g.pspi, g.nref = pspi, 4
    rng = np.random.default_rng(seed)
    d = rng.standard_normal((3, g.nm, g.nt)).astype(np.float32)
    sxs = np.array([sx[0], sx[0] - 150.0, sx[0] + 100.0], np.float32)
    sys_ = np.array([sy[0], sy[0] + 120.0, sy[0] - 60.0], np.float32)
    return g, vel, wav, d, sxs, sys_


@pytest.mark.parametrize("pspi", [False, True], ids=["ssf", "pspi"])
def test_graph_replays_equal_eager_launches(gpu, pspi):
    g, vel, wav, d, sxs, sys_ = case(pspi)
    plan = seiswem.Plan(g, vel, wav)
    dd = seiswem.DeviceArray(d.shape, d)
    plan.set_graphs(False)
    eager = plan.migrate(dd, sxs, sys_).download()
    eager_each = plan.migrate_each(dd, sxs, sys_).download()
    eager_one = plan.migrate(seiswem.DeviceArray((1, g.nm, g.nt), d[1:2]), sxs[1:2], sys_[1:2]).download()
    plan.set_graphs(True)
    for rep in range(4):
        m = plan.migrate(dd, sxs, sys_).download()  # capture on the first, replays after
        assert np.array_equal(m, eager), f"replay {rep} differs from the eager launches"
        # eager work on the same device between replays
        plan.set_graphs(False)
        assert np.array_equal(plan.migrate(dd, sxs, sys_).download(), eager)
        plan.set_graphs(True)
        seiswem.synchronize()
        each = plan.migrate_each(dd, sxs, sys_).download()
        assert np.array_equal(each, eager_each)
        one = plan.migrate(seiswem.DeviceArray((1, g.nm, g.nt), d[1:2]), sxs[1:2], sys_[1:2]).download()
        assert np.array_equal(one, eager_one)
    plan.close()


def test_graph_recaptured_when_chunking_changes(gpu):
    """A plan whose work budget splits the band into frequency chunks, and an accumulated image."""
    g, vel, wav, d, sxs, sys_ = case(False)
    dd = seiswem.DeviceArray(d.shape, d)
    small = seiswem.Plan(g, vel, wav, scratch_bytes=2 << 20)
    small.set_graphs(False)
    ref = small.migrate(dd, sxs, sys_).download()
    small.set_graphs(True)
    m0 = np.random.default_rng(9).standard_normal((g.nm, g.nz)).astype(np.float32)
    for _ in range(3):
        mm = seiswem.DeviceArray(m0.shape, m0)
        got = small.migrate(dd, sxs, sys_, m=mm, accumulate=True).download()
        assert rel_l2(got - m0, ref) <= 1e-6
        assert np.array_equal(small.migrate(dd, sxs, sys_).download(), ref)
    whole = seiswem.Plan(g, vel, wav)
    assert rel_l2(whole.migrate(dd, sxs, sys_).download(), ref) <= 1e-6
    small.close()
    whole.close()


def test_graph_migration_matches_oracle(gpu):
    g, vel, wav, d, sxs, sys_ = case(False, nx=64, ny=50, nz=12)  # Bluestein y axis (nky 100)
    plan = seiswem.Plan(g, vel, wav)
    m = plan.migrate(seiswem.DeviceArray((1, g.nm, g.nt), d[:1]), sxs[:1], sys_[:1]).download()
    m = plan.migrate(seiswem.DeviceArray((1, g.nm, g.nt), d[:1]), sxs[:1], sys_[:1]).download()  # a replay
    p = O.WemParams(nt=g.nt, dt=g.dt, nmx=g.nmx, dmx=g.dmx, nmy=g.nmy, dmy=g.dmy, nz=g.nz, dz=g.dz, fmin=g.fmin,
                    fmax=g.fmax, padt=g.padt, padx=g.padx, pspi=False)
    m_ref = O.wem(p, vel, wav, float(sxs[0]), float(sys_[0]), d=d[0], adj=True)
    err = rel_l2(m, m_ref)
    print(f"graph 3-D migration vs FP64 oracle: rel-L2 {err:.3e}")
    assert err <= 1e-4
    plan.close()
