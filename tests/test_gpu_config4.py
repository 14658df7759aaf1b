"""GPU tests of the serving workload at its own sizes: a seeded subset of the Poisson mix of
migration requests (2-D nx 128..1024, nz 200/400, nt 1000..4000; 3-D 128^2 and 256^2 x 200),
including its largest 2-D request (nx 1024, nz 400, nt 4000) and a 256 x 256 x 200 3-D request.
Each served image equals the plain per-request path (one ShotProfileWEM call, ShotProfileWEM.jl:104)
bit for bit, the largest request gives the same image alone and in the mix, and one 2-D and one
3-D request of the mix agree with the FP64 oracle on a size the oracle covers in seconds."""
import numpy as np
import pytest

import oracle as O
import seiswem
from seiswem import synth
from conftest import rel_l2

pytestmark = pytest.mark.gpu

# request ids of the seeded mix (synth.config4_request, seed 4): 0, 7 are 3-D 256^2, 1 is 3-D
# 128^2, 20 is the largest 2-D shape (1024 x 400, nt 4000), 16 the smallest (128 x 200, nt 1000)
IDS = (0, 1, 2, 3, 4, 5, 6, 7, 15, 16, 20)
LARGEST_2D, SMALLEST_2D, SMALL_3D = 20, 16, 1


@pytest.fixture(scope="module")
def mix():
    records = {}
    reqs = {i: synth.config4_request(i, records=records) for i in IDS}
    g20 = reqs[LARGEST_2D][0]
    assert (g20.nmx, g20.nz, g20.nt, g20.nmy) == (1024, 400, 4000, 1)
    assert (reqs[0][0].nmx, reqs[0][0].nmy, reqs[0][0].nz) == (256, 256, 200)
    assert (reqs[SMALL_3D][0].nmx, reqs[SMALL_3D][0].nmy) == (128, 128)
    return reqs


def plain(req):
    g, vel, wav, d, sx, sy = req
    return seiswem.ShotProfileWEM(d, True, vel=vel, wav=wav, geom=g, sx=sx, sy=sy)


def test_config4_served_equals_plain_and_alone(gpu, mix):
    with seiswem.Server(streams=4, max_inflight=8) as srv:
        ids = {i: srv.submit(*mix[i]) for i in IDS}
        served = {}
        for i in IDS:
            m, st = srv.wait(ids[i])
            assert st["state"] == seiswem.REQ_DONE, st["message"]
            served[i] = m
        stats = srv.stats()
    assert stats["failed"] == 0 and stats["completed"] >= len(IDS)
    for i in IDS:
        print(f"request {i}: {mix[i][0].nmx} x {mix[i][0].nmy} x {mix[i][0].nz}, nt {mix[i][0].nt}", flush=True)
        assert np.array_equal(served[i], plain(mix[i])), f"request {i} differs from the plain path"
        assert np.isfinite(served[i]).all() and np.abs(served[i]).max() > 0
    with seiswem.Server() as srv:
        alone, _ = srv.wait(srv.submit(*mix[LARGEST_2D]))
    assert np.array_equal(alone, served[LARGEST_2D])


def test_config4_2d_request_matches_oracle(gpu, mix):
    g, vel, wav, d, sx, sy = mix[SMALLEST_2D]
    p = O.WemParams(nt=g.nt, dt=g.dt, nmx=g.nmx, dmx=g.dmx, nz=g.nz, dz=g.dz, fmin=g.fmin, fmax=g.fmax, nref=g.nref,
                    padt=g.padt, padx=g.padx, pspi=g.pspi)
    m_ref = O.wem(p, vel, wav, float(sx[0]), d=d[0], adj=True)
    with seiswem.Server() as srv:
        m, _ = srv.wait(srv.submit(g, vel, wav, d, sx, sy))
    err = rel_l2(m, m_ref)
    print(f"2-D request {SMALLEST_2D} vs FP64 oracle: rel-L2 {err:.3e}")
    assert err <= 1e-4


def test_config4_3d_request_matches_oracle(gpu, mix):
    """The 128 x 128 x 200 request's grid, velocity and record on a 1 Hz band (8 bins) and its first
    60 depths: the oracle's literal 3-D transforms finish in seconds there, and the served path
    runs the same kernels as on the full band."""
    import dataclasses
    g, vel, wav, d, sx, sy = mix[SMALL_3D]
    gs = dataclasses.replace(g, fmin=20.0, fmax=21.0, nz=60)
    vs = np.ascontiguousarray(vel[:, :60])
    p = O.WemParams(nt=gs.nt, dt=gs.dt, nmx=gs.nmx, dmx=gs.dmx, nmy=gs.nmy, dmy=gs.dmy, nz=gs.nz, dz=gs.dz,
                    fmin=gs.fmin, fmax=gs.fmax, padt=gs.padt, padx=gs.padx, pspi=False)
    m_ref = O.wem(p, vs, wav, float(sx[0]), float(sy[0]), d=d[0], adj=True)
    with seiswem.Server() as srv:
        m, _ = srv.wait(srv.submit(gs, vs, wav, d, sx, sy))
    err = rel_l2(m, m_ref)
    print(f"3-D request {SMALL_3D} (20-21 Hz, 60 depths) vs FP64 oracle: rel-L2 {err:.3e}")
    assert err <= 1e-4


def test_second_waiter_is_refused(gpu, mix):
    """One thread may wait for a request: a second wait on the same id while the first is blocked
    returns SWEM_EINVAL instead of touching a request the first waiter frees.  With one request in
    flight at a time the probe queues behind a 256 x 256 x 200 request, so the first waiter is
    still blocked when the second call arrives."""
    import ctypes
    import threading
    import time
    lib = seiswem.load()
    with seiswem.Server(max_inflight=1) as srv:
        busy = srv.submit(*mix[0])
        rid = srv.submit(*mix[SMALLEST_2D])
        res = {}
        th = threading.Thread(target=lambda: res.setdefault("m", srv.wait(rid)))
        th.start()
        time.sleep(0.05)
        st = seiswem._lib.RequestStatus()
        rc = lib.swem_batch_wait(srv._s, rid, None, ctypes.byref(st))
        msg = lib.swem_last_error().decode()
        th.join()
        srv.wait(busy)
        assert res["m"][1]["state"] == seiswem.REQ_DONE
    assert rc == seiswem.SWEM_EINVAL and "already has a waiter" in msg, msg
