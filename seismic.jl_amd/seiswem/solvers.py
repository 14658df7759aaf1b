"""Least-squares migration: the reference's solver surface over the GPU operator pair.

  InnerProduct       src/Tools/InnerProduct.jl:1-5      (in-memory: Float32 of real(sum(conj(a) b)))
  CGStep             src/Solvers/CGStep.jl:1-4          m1 = a m1 + b m2
  LinearOperator     src/Operators/LinearOperator.jl:1-21
  DotTest            src/Tools/DotTest.jl:1-12
  ConjugateGradients src/Solvers/ConjugateGradients.jl:1-39  (Algorithm 2 of Scales 1987)

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


def ConjugateGradientsDevice(plan: Plan, d: DeviceArray, sx, sy=None, Niter=10, mu=0.0, tol=1.0e-15):
    """ConjugateGradients.jl:1-39 with m, r, g, s, t resident on the device and double-precision
    inner products.  Returns (m DeviceArray [nm, nz], cost list)."""
    g_ = plan.geom
    cost = []
    r = DeviceArray(d.shape)
    from .wem import copy_device
    copy_device(r, d)
    g = plan.migrate(r, sx, sy)
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
        delta = vec_dot(t, t) + mu * vec_dot(s, s)
        if delta <= tol:
            break
        alpha = gamma / delta
        vec_cgstep(1.0, m, alpha, s)
        vec_cgstep(1.0, r, -alpha, t)
        plan.migrate(r, sx, sy, m=g)
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
