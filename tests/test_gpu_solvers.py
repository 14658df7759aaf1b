"""GPU tests of the least-squares migration path: device vector operations, and the
device-resident CG against the host mirror of the reference solver using the same GPU operator."""
import numpy as np
import pytest

import seiswem
from seiswem import solvers, synth
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def test_vec_dot_and_cgstep(gpu):
    rng = np.random.default_rng(0)
    for n in (1, 1000, 1 << 20):
        x = rng.standard_normal(n).astype(np.float32)
        y = rng.standard_normal(n).astype(np.float32)
        dx, dy = seiswem.DeviceArray(x.shape, x), seiswem.DeviceArray(y.shape, y)
        v = solvers.vec_dot(dx, dy)
        ref = float(np.dot(x.astype(np.float64), y.astype(np.float64)))
        assert abs(v - ref) <= 1e-12 * max(1.0, np.abs(x.astype(np.float64) * y).sum())
        assert solvers.vec_dot(dx, dy) == v  # deterministic
        solvers.vec_cgstep(0.5, dx, -2.0, dy)
        assert np.allclose(dx.download(), np.float32(0.5) * x - np.float32(2.0) * y, rtol=1e-6, atol=1e-6)


def small_lsm():
    g = seiswem.WemGeometry(nt=256, dt=0.004, nmx=64, dmx=10.0, nz=40, dz=10.0, fmin=2.0, fmax=40.0, pspi=False)
    vel = np.repeat((1500.0 + 0.5 * np.arange(g.nz) * g.dz)[None, :], g.nmx, 0).astype(np.float32)
    refl = np.zeros((g.nm, g.nz), np.float32)
    refl[:, 15] = 0.1
    refl[:, 28] = -0.08
    wav = synth.wavelet_trace(g.nt, g.dt, 15.0)
    sx = np.array([120.0, 320.0, 520.0], np.float32)
    return g, vel, refl, wav, sx


def test_device_cg_matches_host_mirror(gpu):
    g, vel, refl, wav, sx = small_lsm()
    op = solvers.WEMOperator(g, vel, wav, sx)
    d_obs = op(refl, False)
    m_host, cost_host = solvers.ConjugateGradients(d_obs, [op], [{}], Niter=5, inner=solvers.InnerProduct64)
    plan = seiswem.Plan(g, vel, wav)
    m_dev, cost_dev = solvers.ConjugateGradientsDevice(plan, seiswem.DeviceArray(d_obs.shape, d_obs), sx, Niter=5)
    # the host mirror keeps r, s, g in double (numpy promotes the float32 records), the device in
    # float32: the normalised misfits agree to ~1e-5 after cancellation in r = r - alpha t
    assert rel_l2(m_dev.download(), m_host) <= 1e-5
    assert np.allclose(cost_dev, cost_host, rtol=1e-4)
    assert all(np.diff(cost_dev) < 0)  # the misfit decreases every iteration


def test_wem_operator_dot_test(gpu):
    g, vel, refl, wav, sx = small_lsm()
    op = solvers.WEMOperator(g, vel, wav, sx)
    rng = np.random.default_rng(3)
    m = rng.standard_normal((g.nm, g.nz)).astype(np.float32)
    d = rng.standard_normal((sx.size, g.nm, g.nt)).astype(np.float32)
    a, b = solvers.DotTest(m, d, [op], [{}], inner=solvers.InnerProduct64)
    assert abs(a - b) / abs(b) <= 1e-5


def test_vec_mul(gpu):
    rng = np.random.default_rng(6)
    for n in (1, 777, 1 << 20):
        x = rng.standard_normal(n).astype(np.float32)
        w = rng.standard_normal(n).astype(np.float32)
        dx, dw, dy = seiswem.DeviceArray(x.shape, x), seiswem.DeviceArray(w.shape, w), seiswem.DeviceArray(x.shape)
        solvers.vec_mul(dx, dw, dy)
        assert np.array_equal(dy.download(), x * w)
        solvers.vec_mul(dx, dw)  # in place
        assert np.array_equal(dx.download(), x * w)


def test_lswem_matches_host_weighted_chain(gpu):
    """ShotProfileLSWEM (data weights from CalculateSampling with dead traces, a mute, the GPU WEM
    pair, all on the device) against the host CG over the chain [WeightingOp, WEMOperator] with
    the same weight vector, and the weighted chain's dot test."""
    g, vel, refl, wav, sx = small_lsm()
    op = solvers.WEMOperator(g, vel, wav, sx)
    d_obs = op(refl, False)
    d_obs[0, 5:9] = 0.0    # dead traces
    d_obs[2, 40] = 0.0
    kw = dict(tmute=0.05, vmute=1500.0, taper=0.04)
    plan = seiswem.Plan(g, vel, wav)
    m_dev, cost_dev = solvers.ShotProfileLSWEM(d_obs, plan, sx, Niter=5, **kw)
    wd, _ = solvers.CalculateSampling(d_obs, h=True)
    assert wd[0, 5:9].sum() == 0 and wd[1].min() == 1
    gx = g.omx + g.dmx * np.arange(g.nm)
    w = np.stack([wd[i] * solvers.SeisMute(np.ones((g.nm, g.nt), np.float32), offset=np.abs(gx - sx[i]),
                                           dt=g.dt, **kw) for i in range(sx.size)])
    ops, pars = [solvers.WeightingOp, op], [dict(w=w), {}]
    m_host, cost_host = solvers.ConjugateGradients(d_obs, ops, pars, Niter=5, inner=solvers.InnerProduct64)
    err = rel_l2(m_dev, m_host)
    print(f"LSWEM device vs host chain: rel-L2 {err:.3e}")
    assert err <= 1e-5
    assert np.allclose(cost_dev, cost_host, rtol=1e-4)
    rng = np.random.default_rng(8)
    a, b = solvers.DotTest(rng.standard_normal((g.nm, g.nz)).astype(np.float32),
                           rng.standard_normal(d_obs.shape).astype(np.float32), ops, pars, inner=solvers.InnerProduct64)
    assert abs(a - b) / abs(b) <= 1e-5
