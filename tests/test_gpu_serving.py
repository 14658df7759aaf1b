"""GPU tests of the serving loop (swem_server_* / swem_batch_*): served results equal the plain
per-request path bit for bit, a request gives the same image alone and in a mix, admission
respects the device arena, and an oversized request fails cleanly."""
import numpy as np
import pytest

import oracle as O
import seiswem
from seiswem import synth
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def mixed(n, seed=0):
    """Small requests of every kind: 2-D SSF, 2-D PSPI (compat and linear; nx 100 takes the
    Bluestein kernels), 3-D SSF (every second one on 50 x 40 points: Bluestein lengths 100 x 80)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        kind = i % 4
        if kind == 3:
            nx3, ny3 = (64, 64) if (i // 4) % 2 == 0 else (50, 40)
            g, vel, _, wav, sx, sy = synth.config3(seed=i, nx=nx3, ny=ny3, nz=16, nt=128, fmax=20.0)
            mode = seiswem.PSPI_COMPAT
        else:
            nx = int(rng.choice([64, 100, 128]))
            g = seiswem.WemGeometry(nt=256, dt=0.004, nmx=nx, dmx=10.0, nz=30, dz=10.0, fmin=2.0, fmax=40.0, nref=4,
                                    pspi=kind > 0)
            vel = synth.blob_velocity(nx, g.nz, g.dmx, g.dz, rng, nblobs=3, v0=2000.0)
            wav = synth.wavelet_trace(g.nt, g.dt, 15.0)
            sx, sy = np.array([nx * 5.0], np.float32), None
            mode = seiswem.PSPI_LINEAR if kind == 2 else seiswem.PSPI_COMPAT
        d = rng.standard_normal((1, g.nm, g.nt)).astype(np.float32)
        out.append((g, vel, wav, d, sx, sy, mode))
    return out


def plain(req):
    g, vel, wav, d, sx, sy, mode = req
    return seiswem.ShotProfileWEM(d, True, vel=vel, wav=wav, geom=g, sx=sx, sy=sy, pspi_mode=mode)


def test_served_equals_plain_path():
    reqs = mixed(8)
    with seiswem.Server(streams=4, max_inflight=8) as srv:
        ids = [srv.submit(g, vel, wav, d, sx, sy, pspi_mode=mode) for g, vel, wav, d, sx, sy, mode in reqs]
        for rid, req in zip(ids, reqs):
            m, st = srv.wait(rid)
            assert st["state"] == seiswem.REQ_DONE
            assert np.array_equal(m, plain(req)), f"request {rid} differs from the plain path"


def test_same_result_alone_and_in_mix():
    reqs = mixed(13, seed=3)
    probe = reqs[6]
    with seiswem.Server() as srv:
        alone, _ = srv.wait(srv.submit(*probe[:6], pspi_mode=probe[6]))
    with seiswem.Server(streams=4, max_inflight=8) as srv:
        ids = [srv.submit(*r[:6], pspi_mode=r[6]) for r in reqs]
        results = [srv.wait(i) for i in ids]
        st = srv.stats()
    mix = results[6][0]
    assert np.array_equal(alone, mix)
    # the requests overlapped: some request was admitted before an earlier one finished
    spans = sorted((s["admit_ms"], s["done_ms"]) for _, s in results)
    assert any(spans[i + 1][0] < spans[i][1] for i in range(len(spans) - 1))
    assert st["completed"] >= 13 and st["failed"] == 0


def test_served_matches_oracle():
    g, vel, wav, d, sx, sy, mode = mixed(2, seed=5)[1]
    p = O.WemParams(nt=g.nt, dt=g.dt, nmx=g.nmx, dmx=g.dmx, nz=g.nz, dz=g.dz, fmin=g.fmin, fmax=g.fmax, nref=g.nref,
                    padt=g.padt, padx=g.padx, pspi=g.pspi, pspi_mode=mode)
    m_ref = O.wem(p, vel, wav, float(sx[0]), d=d[0], adj=True)
    with seiswem.Server() as srv:
        m, _ = srv.wait(srv.submit(g, vel, wav, d, sx, sy, pspi_mode=mode))
    assert rel_l2(m, m_ref) <= 1e-4


def test_admission_respects_arena_and_rejects_oversized():
    reqs = mixed(6, seed=9)[:3]
    # records alone (256 x 256 traces x 1000 samples = 262 MB) exceed the arena
    big_g = seiswem.WemGeometry(nt=1000, dt=0.004, nmx=256, dmx=12.5, nmy=256, dmy=12.5, nz=50, dz=10.0, pspi=False)
    arena = 96 << 20
    with seiswem.Server(device_bytes=arena, pinned_bytes=64 << 20) as srv:
        ids = [srv.submit(*r[:6], pspi_mode=r[6]) for r in reqs * 3]
        for i in ids:
            m, st = srv.wait(i)
            assert st["state"] == seiswem.REQ_DONE
        vel = np.full((big_g.nm, big_g.nz), 2000.0, np.float32)
        bid = srv.submit(big_g, vel, np.zeros(big_g.nt, np.float32), np.zeros((1, big_g.nm, big_g.nt), np.float32),
                         [1600.0], [1600.0])
        with pytest.raises(seiswem.SwemError) as e:
            srv.wait(bid)
        assert e.value.code == seiswem.SWEM_ENOMEM
        st = srv.stats()
        assert 0 < st["device_bytes_peak"] <= arena
