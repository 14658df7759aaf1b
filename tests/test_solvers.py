"""CPU tests of the solver mirror (seiswem.solvers): ConjugateGradients against the closed-form
damped least-squares solution on a dense matrix (as test/test_CG.jl:1-23 of the reference), the
operator chaining of LinearOperator, DotTest, and the Float32 rounding of InnerProduct."""
import numpy as np

from seiswem import solvers


def matrix_op(x, adj, matrix):
    return matrix.T @ x if adj else matrix @ x


def test_cg_matches_closed_form_damped_least_squares():
    rng = np.random.default_rng(0)
    nd, nm = 1000, 300
    L = rng.standard_normal((nd, nm))
    m = rng.standard_normal(nm)
    d = L @ m + 0.2 * rng.standard_normal(nd)
    mu = 0.2
    m1 = np.linalg.solve(L.T @ L + mu * np.eye(nm), L.T @ d)
    m2, cost = solvers.ConjugateGradients(d, [matrix_op], [dict(matrix=L)], Niter=nm, mu=mu,
                                          inner=solvers.InnerProduct64)
    quality = 10 * np.log10(np.linalg.norm(m1) / np.linalg.norm(m2 - m1))
    assert quality > 50.0
    assert cost[0] == 1.0 and all(np.diff(cost) <= 1e-12)


def test_cg_reference_float32_inner_product():
    """With the reference's Float32 InnerProduct the iteration still converges (quality > 10 dB,
    the reference's own bar)."""
    rng = np.random.default_rng(1)
    L = rng.standard_normal((200, 60))
    d = L @ rng.standard_normal(60)
    m1 = np.linalg.solve(L.T @ L + 0.1 * np.eye(60), L.T @ d)
    m2, _ = solvers.ConjugateGradients(d, [matrix_op], [dict(matrix=L)], Niter=60, mu=0.1)
    assert 10 * np.log10(np.linalg.norm(m1) / np.linalg.norm(m2 - m1)) > 10.0


def test_linear_operator_chain_order_and_dot_test():
    rng = np.random.default_rng(2)
    A = rng.standard_normal((40, 30))
    B = rng.standard_normal((30, 20))
    ops, pars = [matrix_op, matrix_op], [dict(matrix=A), dict(matrix=B)]
    # forward applies the chain in reverse order: d = A (B m); adjoint: m = B^T (A^T d)
    x = rng.standard_normal(20)
    assert np.allclose(solvers.LinearOperator(x, ops, pars, adj=False), A @ (B @ x))
    y = rng.standard_normal(40)
    assert np.allclose(solvers.LinearOperator(y, ops, pars, adj=True), B.T @ (A.T @ y))
    a, b = solvers.DotTest(x, y, ops, pars, inner=solvers.InnerProduct64)
    assert abs(a - b) <= 1e-12 * abs(a)


def test_inner_product_is_float32_of_real_part():
    a = np.array([1 + 2j, 3 - 1j])
    b = np.array([2 - 1j, 1 + 1j])
    v = solvers.InnerProduct(a, b)
    assert isinstance(v, np.float32)
    assert v == np.float32(np.real(np.sum(np.conj(a) * b)))
    assert solvers.CGStep(np.ones(3), np.full(3, 2.0), a=0.5, b=-1.0).tolist() == [-1.5, -1.5, -1.5]


def test_seis_mute_closed_form():
    """SeisMute.jl:1-19: t = sqrt(tmute^2 + (offset/vmute)^2); samples at it*dt <= t - taper are
    zeroed, those in (t - taper, t] scaled by 1 - (t - it*dt)/taper, later ones untouched."""
    dt, nt = 0.01, 60
    x = np.ones((3, nt), np.float32)
    off = np.array([0.0, 300.0, 800.0])
    out = solvers.SeisMute(x, offset=off, tmute=0.2, vmute=1000.0, taper=0.1, dt=dt)
    for ix, o in enumerate(off):
        t = np.hypot(0.2, o / 1000.0)
        for it in range(nt):
            ts = it * dt
            want = (0.0 if ts <= t - 0.1 else 1.0 - (t - ts) / 0.1) if ts <= t else 1.0
            assert abs(out[ix, it] - want) <= 1e-6, (ix, it)
    # offset 0, tmute 0.2: samples 0..10 zeroed (ts <= 0.1), ramp on 11..20, 1 from 21 on
    assert out[0, 10] == 0.0 and out[0, 21] == 1.0 and abs(out[0, 15] - 0.5) < 1e-6


def test_calculate_sampling_live_and_dead_traces():
    d = np.zeros((2, 4, 8), np.float32)
    d[0, 1, 3] = 1.0
    d[1, 2, 0] = 1e-6          # sum of squares 1e-12 <= 1e-10, sqrt 1e-6 > 1e-10
    wd = solvers.CalculateSampling(d)
    assert wd.shape == d.shape and wd.dtype == np.float32
    assert np.array_equal(wd[:, :, 0], np.array([[0, 1, 0, 0], [0, 0, 0, 0]], np.float32))
    wd2, h = solvers.CalculateSampling(d, h=True)
    assert h is True
    assert np.array_equal(wd2[:, :, 5], np.array([[0, 1, 0, 0], [0, 0, 1, 0]], np.float32))


def test_weighted_chain_cg_equals_cg_on_weighted_matrix():
    """[WeightingOp, SeisMute-as-operator, L] through LinearOperator is diag(w) diag(mute) L: CG on
    the chain reproduces CG on the explicitly weighted matrix, and the chain passes the dot test."""
    rng = np.random.default_rng(4)
    ntr, nt, nm = 6, 20, 40
    L = rng.standard_normal((ntr * nt, nm))
    w = (rng.random((ntr, nt)) > 0.2).astype(np.float64)
    off = np.linspace(0, 500, ntr)

    def mute_op(x, adj, **kw):
        return solvers.SeisMute(x.reshape(ntr, nt), **kw).reshape(x.shape)

    mute_kw = dict(offset=off, tmute=0.05, vmute=2000.0, taper=0.05, dt=0.01)
    mvec = solvers.SeisMute(np.ones((ntr, nt)), **mute_kw).ravel()
    ops = [solvers.WeightingOp, mute_op, matrix_op]
    pars = [dict(w=w.ravel()), mute_kw, dict(matrix=L)]
    d = L @ rng.standard_normal(nm)
    m1, c1 = solvers.ConjugateGradients(d, ops, pars, Niter=15, inner=solvers.InnerProduct64)
    Lw = (w.ravel() * mvec)[:, None] * L
    m2, c2 = solvers.ConjugateGradients(d, [matrix_op], [dict(matrix=Lw)], Niter=15, inner=solvers.InnerProduct64)
    assert np.allclose(m1, m2, rtol=1e-10, atol=1e-12) and np.allclose(c1, c2, rtol=1e-10)
    a, b = solvers.DotTest(rng.standard_normal(nm), rng.standard_normal(ntr * nt), ops, pars,
                           inner=solvers.InnerProduct64)
    assert abs(a - b) <= 1e-10 * abs(b)
