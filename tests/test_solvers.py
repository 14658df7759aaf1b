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
