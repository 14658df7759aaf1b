"""Least-squares migration: the reference's solver surface over the GPU operator pair.

  InnerProduct       src/Tools/InnerProduct.jl:1-5      (in-memory: Float32 of real(sum(conj(a) b)))
  CGStep             src/Solvers/CGStep.jl:1-4          m1 = a m1 + b m2
  LinearOperator     src/Operators/LinearOperator.jl:1-21
  DotTest            src/Tools/DotTest.jl:1-12
  ConjugateGradients src/Solvers/ConjugateGradients.jl:1-39  (Algorithm 2 of Scales 1987)
  WeightingOp        src/Operators/WeightingOp.jl:1-5
  SeisMute           src/Processing/SeisMute.jl:1-19
  CalculateSampling  src/Tools/CalculateSampling.jl:1-40
  ShotProfileLSWEM   src/Imaging/ShotProfileLSWEM.jl:1-21 (data weights, mute, WEM; repaired, see below)

Operators follow the reference convention op(x, adj; params...) -> y.  `WEMOperator` is the
shot-profile migration / demigration pair on the device (ShotProfileWEM with nangx = nangy = 1).
`ConjugateGradientsDevice` runs the same iteration with model, data and residual resident in HBM
(vector updates and FP64-accumulated inner products on the device through swem_vec_*).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check, load
from .wem import DeviceArray, Plan, WemGeometry


def InnerProduct(in1, in2) -> np.float32:
    """In-memory InnerProduct.jl:1-5: Float32 of real(sum(conj(in1) .* in2))."""
    return np.float32(np.real(np.sum(np.conj(np.ravel(in1)) * np.ravel(in2))))


def InnerProduct64(in1, in2) -> float:
    """The same sum accumulated and returned in double precision."""
    return float(np.real(np.sum(np.conj(np.ravel(in1)).astype(np.complex128) * np.ravel(in2))))


def CGStep(m1, m2, a=1.0, b=1.0):
    """CGStep.jl:1-4: a*m1 + b*m2."""
    return a * m1 + b * m2


def LinearOperator(x, operators, parameters, adj=True):
    """LinearOperator.jl:1-21: chain of operators, applied in order for the adjoint and in
    reverse order for the forward."""
    if adj:
        d = np.copy(x)
        m = None
        for op, par in zip(operators, parameters):
            m = op(d, True, **par)
            d = np.copy(m)
        return m
    m = np.copy(x)
    d = None
    for op, par in zip(reversed(operators), reversed(parameters)):
        d = op(m, False, **par)
        m = np.copy(d)
    return d


def WeightingOp(in_, adj=True, w=1.0, **_):
    """WeightingOp.jl:1-5: in .* w (a diagonal operator: the adjoint is the same product)."""
    return np.asarray(in_) * w


def SeisMute(in_, offset=0.0, tmute=0.0, vmute=1500.0, taper=0.1, dt=0.001):
    """SeisMute.jl:1-19 on traces of nt samples (numpy [..., nx, nt], time fastest: the memory
    of Julia's (nt, nx)).  A sample at time t_s = it dt with t_s <= t = sqrt(tmute^2 +
    (offset/vmute)^2) is zeroed, except on the last `taper` seconds before t where it is scaled
    by 1 - (t - t_s) / taper.  `offset` is one value per trace (or a scalar)."""
    x = np.array(in_, copy=True)
    nt, nx = x.shape[-1], x.shape[-2]
    off = np.broadcast_to(np.asarray(offset, np.float64), (nx,))
    t = np.sqrt(tmute ** 2 + (off / vmute) ** 2)[:, None]
    ts = np.arange(nt, dtype=np.float64)[None, :] * dt
    factor = np.where(ts <= t, np.where(ts > t - taper, 1.0 - (t - ts) / taper, 0.0), 1.0)
    return (x * factor.astype(x.dtype)).astype(x.dtype)


def CalculateSampling(in_, h=None, cutoff=1e-10):
    """CalculateSampling.jl: 1 on live traces, 0 on dead ones, per trace of nt samples
    ([..., nt]).  Without headers a trace is live when sum(x^2) > cutoff (lines 1-18); with
    headers, as the file form used by ShotProfileLSWEM runs it, when sqrt(sum(x^2)) > cutoff
    (lines 20-40), and (wd, h) is returned."""
    x = np.asarray(in_, np.float64)
    e = np.sum(x * x, axis=-1)
    live = (np.sqrt(e) if h is not None else e) > cutoff
    wd = np.broadcast_to(live[..., None], x.shape).astype(np.float32)
    return (wd, h) if h is not None else wd


def DotTest(m_rand, d_rand, operators, parameters, inner=InnerProduct):
    """DotTest.jl:1-12: (<d_rand, L m_rand>, <m_rand, L^H d_rand>)."""
    m_adj = LinearOperator(d_rand, operators, parameters, adj=True)
    d_fwd = LinearOperator(m_rand, operators, parameters, adj=False)
    return inner(d_rand, d_fwd), inner(m_rand, m_adj)


def ConjugateGradients(d, operators, parameters, Niter=10, mu=0.0, tol=1.0e-15, inner=InnerProduct):
    """ConjugateGradients.jl:1-39 on in-memory arrays; returns (m, cost).  `inner` defaults to
    the reference's Float32 InnerProduct."""
    cost = []
    r = np.copy(d)
    g = LinearOperator(r, operators, parameters, adj=True)
    m = np.zeros_like(g)
    s = np.copy(g)
    gamma = inner(g, g)
    gamma00 = gamma
    cost0 = inner(r, r)
    cost.append(1.0)
    for _ in range(Niter):
        t = LinearOperator(s, operators, parameters, adj=False)
        delta = inner(t, t) + mu * inner(s, s)
        if delta <= tol:
            break
        alpha = gamma / delta
        m = m + alpha * s
        r = r - alpha * t
        g = LinearOperator(r, operators, parameters, adj=True)
        g = g - mu * m
        gamma0 = gamma
        gamma = inner(g, g)
        cost1 = inner(r, r) + mu * inner(m, m)
        cost.append(float(cost1 / cost0))
        beta = gamma / gamma0
        s = beta * s + g
        if np.sqrt(gamma) <= np.sqrt(gamma00) * tol:
            break
    return m, cost


class WEMOperator:
    """Shot-profile WEM as an operator op(x, adj): adj maps records [nshot, nm, nt] to the stacked
    image [nm, nz]; forward maps an image to records.  Host arrays in and out; the plan (velocity
    tables, source spectrum) stays on the device between calls."""

    def __init__(self, geom: WemGeometry, vel, wav, sx, sy=None, **plan_kw):
        self.geom = geom
        self.sx = np.ascontiguousarray(np.atleast_1d(sx), dtype=np.float32)
        self.sy = None if sy is None else np.ascontiguousarray(np.atleast_1d(sy), dtype=np.float32)
        self.plan = Plan(geom, vel, wav, **plan_kw)

    def __call__(self, x, adj, **_):
        g = self.geom
        if adj:
            d = DeviceArray((self.sx.size, g.nm, g.nt), np.asarray(x, np.float32))
            return self.plan.migrate(d, self.sx, self.sy).download()
        m = DeviceArray((g.nm, g.nz), np.asarray(x, np.float32))
        return self.plan.model(m, self.sx, self.sy).download()


# ------------------------------------------------------------------ device-resident solver
def vec_dot(x: DeviceArray, y: DeviceArray) -> float:
    """<x, y> of two device vectors, accumulated in double in a fixed order."""
    out = ctypes.c_double()
    check(load().swem_vec_dot(x.ptr, y.ptr, x.nbytes // 4, ctypes.byref(out), None))
    return out.value


def vec_cgstep(a: float, m1: DeviceArray, b: float, m2: DeviceArray):
    """m1 = a m1 + b m2 on the device (CGStep.jl:1-4)."""
    check(load().swem_vec_cgstep(a, m1.ptr, b, m2.ptr, m1.nbytes // 4, None))


def vec_mul(x: DeviceArray, w: DeviceArray, y: DeviceArray | None = None) -> DeviceArray:
    """y = x .* w on the device (WeightingOp.jl:1-5); in place when y is None."""
    if w.nbytes != x.nbytes:
        raise ValueError(f"vec_mul: weights of {w.nbytes} bytes for a vector of {x.nbytes}")
    y = x if y is None else y
    check(load().swem_vec_mul(x.ptr, w.ptr, y.ptr, x.nbytes // 4, None))
    return y


def ConjugateGradientsDevice(plan: Plan, d: DeviceArray, sx, sy=None, Niter=10, mu=0.0, tol=1.0e-15,
                             weights: DeviceArray | None = None):
    """ConjugateGradients.jl:1-39 with m, r, g, s, t resident on the device and double-precision
    inner products.  With `weights` (a device vector shaped like d) the operator is the chain
    [WeightingOp, ShotProfileWEM] of LinearOperator.jl: L = diag(w) WEM, L^H = WEM^H diag(w).
    Returns (m DeviceArray [nm, nz], cost list)."""
    g_ = plan.geom
    cost = []
    r = DeviceArray(d.shape)
    from .wem import copy_device
    copy_device(r, d)
    wr = DeviceArray(d.shape) if weights is not None else None

    def adjoint(x, out=None):
        if weights is not None:
            x = vec_mul(x, weights, wr)
        return plan.migrate(x, sx, sy, m=out)

    g = adjoint(r)
    m = DeviceArray((g_.nm, g_.nz))
    m.zero()
    s = DeviceArray(g.shape)
    copy_device(s, g)
    t = DeviceArray(d.shape)
    gamma = vec_dot(g, g)
    gamma00 = gamma
    cost0 = vec_dot(r, r)
    cost.append(1.0)
    for _ in range(Niter):
        plan.model(s, sx, sy, d=t)
        if weights is not None:
            vec_mul(t, weights)
        delta = vec_dot(t, t) + mu * vec_dot(s, s)
        if delta <= tol:
            break
        alpha = gamma / delta
        vec_cgstep(1.0, m, alpha, s)
        vec_cgstep(1.0, r, -alpha, t)
        adjoint(r, g)
        if mu:
            vec_cgstep(1.0, g, -mu, m)
        gamma0 = gamma
        gamma = vec_dot(g, g)
        cost1 = vec_dot(r, r) + mu * vec_dot(m, m)
        cost.append(cost1 / cost0)
        beta = gamma / gamma0
        vec_cgstep(beta, s, 1.0, g)
        if np.sqrt(gamma) <= np.sqrt(gamma00) * tol:
            break
    return m, cost


def ShotProfileLSWEM(d, plan: Plan, sx, sy=None, *, Niter=10, mu=0.0, tol=1.0e-15, wd=None, tmute=0.0,
                     vmute=999999.0, taper=0.1, offset=None, precon=False):
    """Least-squares shot-profile migration, ShotProfileLSWEM.jl:1-21 repaired.

    The reference chain is [ApplyDataWeights, SeisMute, ShotProfileWEM] under the file-based
    ConjugateGradients with the sampling operator of CalculateSampling(d) as data weights,
    tmute = 0 and vmute = 999999; as written it does not run (ApplyDataWeights is not defined
    anywhere, ConjugateGradients has no (m, m0, d, cost, param) method, and SeisMute reads the
    header `h` that shotwem never writes).  Here the chain is WeightingOp (the defined data-weight
    operator) with wd = CalculateSampling(d, h) (the rule of the file form the reference calls),
    then SeisMute with per-trace offsets (default: source-receiver distance from the geometry),
    then the GPU WEM pair; both weights are diagonal, so they are applied as one device weight
    vector inside ConjugateGradientsDevice.  precon (smoothing preconditioners
    SeisSmoothStructure / SeisSmoothGathers) is not part of this build.

    d: records [nshot, nm, nt] (host array or DeviceArray).  Returns (m host [nm, nz], cost)."""
    if precon:
        raise NotImplementedError("ShotProfileLSWEM: the smoothing preconditioner is not part of this build")
    g = plan.geom
    sx = np.ascontiguousarray(np.atleast_1d(sx), dtype=np.float32)
    nshot = sx.size
    d_dev = d if isinstance(d, DeviceArray) else DeviceArray((nshot, g.nm, g.nt), np.asarray(d, np.float32))
    if wd is None:
        d_host = d_dev.download().reshape(nshot, g.nm, g.nt)
        wd, _ = CalculateSampling(d_host, h=True)
    wd = np.broadcast_to(np.asarray(wd, np.float32), (nshot, g.nm, g.nt))
    if offset is None:
        gx = g.omx + g.dmx * (np.arange(g.nm) // g.nmy)
        gy = g.omy + (g.dmy if g.dmy > 0 else g.dmx) * (np.arange(g.nm) % g.nmy)  # shotwem.c:68: dy = dx
        syv = np.zeros(nshot) if sy is None else np.atleast_1d(np.asarray(sy, np.float64))
        offset = np.sqrt((gx[None, :] - sx[:, None]) ** 2 + (gy[None, :] - syv[:, None]) ** 2)
    offset = np.broadcast_to(np.asarray(offset, np.float64), (nshot, g.nm))
    w = np.empty((nshot, g.nm, g.nt), np.float32)
    ones = np.ones((g.nm, g.nt), np.float32)
    for i in range(nshot):
        w[i] = wd[i] * SeisMute(ones, offset=offset[i], tmute=tmute, vmute=vmute, taper=taper, dt=g.dt)
    w_dev = DeviceArray(w.shape, w)
    m, cost = ConjugateGradientsDevice(plan, d_dev, sx, sy, Niter=Niter, mu=mu, tol=tol, weights=w_dev)
    return m.download(), cost
