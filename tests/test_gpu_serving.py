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


def survey(n, three_d=False, seed=11, vel_seed=None):
    """n single-shot requests over one velocity model and wavelet (the shots of one survey), each
    with its own records and source position: 2-D PSPI 128 x 40, or 3-D SSF 64 x 64 x 12."""
    rng = np.random.default_rng(seed)
    vrng = np.random.default_rng(seed if vel_seed is None else vel_seed)
    if three_d:
        g, vel, _, wav, _, _ = synth.config3(seed=seed if vel_seed is None else vel_seed, nx=64, ny=64, nz=12,
                                             nt=128, fmax=20.0)
        xs, ys = rng.uniform(100.0, 700.0, n), rng.uniform(100.0, 700.0, n)
    else:
        g = seiswem.WemGeometry(nt=256, dt=0.004, nmx=128, dmx=10.0, nz=40, dz=10.0, fmin=2.0, fmax=40.0, nref=4,
                                pspi=True)
        vel = synth.blob_velocity(g.nmx, g.nz, g.dmx, g.dz, vrng, nblobs=3, v0=2000.0)
        wav = synth.wavelet_trace(g.nt, g.dt, 15.0)
        xs, ys = rng.uniform(50.0, 1200.0, n), None
    reqs = []
    for i in range(n):
        d = rng.standard_normal((1, g.nm, g.nt)).astype(np.float32)
        sy = None if ys is None else np.array([ys[i]], np.float32)
        reqs.append((g, vel, wav, d, np.array([xs[i]], np.float32), sy, seiswem.PSPI_COMPAT))
    return reqs


def gate():
    """A 3-D request that keeps a one-slot server busy while the requests behind it queue."""
    g, vel, _, wav, sx, sy = synth.config3(seed=1, nx=128, ny=128, nz=100, nt=512, fmax=40.0)
    d = np.random.default_rng(2).standard_normal((1, g.nm, g.nt)).astype(np.float32)
    return (g, vel, wav, d, sx, sy, seiswem.PSPI_COMPAT)


def serve_behind_gate(reqs, **cfg):
    """Submits the gate, waits until it runs, queues reqs behind it; returns images and stats."""
    with seiswem.Server(max_inflight=1, **cfg) as srv:
        gid = srv.submit(*gate()[:6])
        while srv.poll(gid)["state"] == seiswem.REQ_PENDING:
            pass
        ids = [srv.submit(*r[:6], pspi_mode=r[6]) for r in reqs]
        srv.wait(gid)
        out = [srv.wait(i) for i in ids]
        st = srv.stats()
    assert all(s["state"] == seiswem.REQ_DONE for _, s in out)
    return [m for m, _ in out], st


@pytest.mark.parametrize("three_d", [False, True], ids=["2d_pspi", "3d_ssf"])
def test_fused_requests_equal_alone(three_d):
    """Ten single-shot requests of one survey queued together run as one launch of ten shots on one
    plan (fusion); each image equals the request served alone and the plain path, bit for bit."""
    reqs = survey(10, three_d)
    fused, st = serve_behind_gate(reqs, max_fuse=32)
    alone, st1 = serve_behind_gate(reqs, max_fuse=1)
    print(f"fusion ({'3-D' if three_d else '2-D'}): {st['fused_launches']} launch(es) carried "
          f"{st['fused_requests']} requests", flush=True)
    assert st["fused_launches"] == 1 and st["fused_requests"] == 10
    assert st1["fused_launches"] == 0
    for i, r in enumerate(reqs):
        assert np.array_equal(fused[i], alone[i]), f"request {i}: fused image differs from alone"
        if i % 3 == 0:
            assert np.array_equal(fused[i], plain(r)), f"request {i}: fused image differs from the plain path"
    if not three_d:
        g, vel, wav, d, sx, sy, mode = reqs[4]
        p = O.WemParams(nt=g.nt, dt=g.dt, nmx=g.nmx, dmx=g.dmx, nz=g.nz, dz=g.dz, fmin=g.fmin, fmax=g.fmax,
                        nref=g.nref, padt=g.padt, padx=g.padx, pspi=g.pspi, pspi_mode=mode)
        assert rel_l2(fused[4], O.wem(p, vel, wav, float(sx[0]), d=d[0], adj=True)) <= 1e-4


def test_fusion_joins_only_equal_requests_and_respects_max_fuse():
    """Behind the gate: six requests of survey A, three of survey B (another velocity), a two-shot
    request and a request with another wavelet, both on A's velocity.  With max_fuse 4: A runs as
    launches of 4 and 2, B as one of 3, the other two alone; every image equals max_fuse 1."""
    a = survey(6, seed=21)
    b = survey(3, seed=22, vel_seed=99)
    g, vel, wav, d, sx, sy, mode = a[0]
    two = (g, vel, wav, np.concatenate([a[1][3], a[2][3]]), np.concatenate([a[1][4], a[2][4]]), None, mode)
    other_wav = (g, vel, synth.wavelet_trace(g.nt, g.dt, 25.0), a[3][3], a[3][4], None, mode)
    reqs = [a[0], b[0], a[1], two, a[2], b[1], other_wav, a[3], a[4], b[2], a[5]]
    fused, st = serve_behind_gate(reqs, max_fuse=4)
    alone, _ = serve_behind_gate(reqs, max_fuse=1)
    print(f"fusion of the mix: {st['fused_launches']} launches carried {st['fused_requests']} requests", flush=True)
    assert st["fused_launches"] == 3 and st["fused_requests"] == 9
    for i in range(len(reqs)):
        assert np.array_equal(fused[i], alone[i]), f"request {i} differs"


def test_server_rejects_bad_max_fuse():
    with pytest.raises(seiswem.SwemError) as e:
        seiswem.Server(max_fuse=0)
    assert e.value.code == seiswem.SWEM_EINVAL
