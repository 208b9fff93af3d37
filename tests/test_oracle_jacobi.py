"""The oracle's block Jacobi (Dune::SeqJac(A, PreconditionerOrder, w) behind
PreconditionerWrapperJacobi, istlpreconditionerwrappers.hh:84-89): `order`
sweeps x <- (1 - w) x + w D^{-1} (d - (L + U) x) from the output vector's
content.  Pinned by a dense restatement of the sweeps, the fixed point of
many sweeps (the solution of the system) and the reference default order 0,
with which the first BiCGStab iteration breaks down."""
import numpy as np
import pytest

import oracle as orc
from cases import case, mixed_state


def _dense(N, rp, ci, vals):
    A = np.zeros((3 * N, 3 * N))
    B = vals.reshape(-1, 3, 3)
    for i in range(N):
        for k in range(rp[i], rp[i + 1]):
            A[3 * i:3 * i + 3, 3 * ci[k]:3 * ci[k] + 3] = B[k]
    return A


def _system(dims=(4, 3, 2), seed=3, dominant=False):
    c = case(*dims, seed=seed)
    prob = orc.problem_from_case(c)
    pv, mean = mixed_state(c, seed=seed)
    J, r = orc.linearize(prob, pv, mean, orc.storage(prob, *mixed_state(c, seed=seed + 1)), 43200.0)
    pat = orc.pattern(prob)
    if dominant:
        # a synthetic matrix on the same pattern whose diagonal blocks dominate
        rng = np.random.default_rng(seed)
        J = rng.uniform(-1.0, 1.0, J.size)
        B = J.reshape(-1, 3, 3)
        for i in range(c.N):
            B[pat["diag_off"][i]] += 40.0 * np.eye(3)
    return c.N, pat, J, r


@pytest.mark.parametrize("order,relax", [(1, 1.0), (2, 1.0), (3, 0.7), (5, 1.3)])
def test_jacobi_sweeps_match_a_dense_restatement(order, relax):
    N, pat, J, r = _system()
    rp, ci = pat["row_ptr"], pat["col_idx"]
    A = _dense(N, rp, ci, J)
    D = np.zeros_like(A)
    for i in range(N):
        D[3 * i:3 * i + 3, 3 * i:3 * i + 3] = A[3 * i:3 * i + 3, 3 * i:3 * i + 3]
    Dinv = np.linalg.inv(D)
    v0 = np.random.default_rng(1).standard_normal(3 * N)
    x = v0.copy()
    for _ in range(order):
        x = (1.0 - relax) * x + relax * (Dinv @ (r - (A - D) @ x))
    rc, v = orc.jacobi_apply(N, 3, rp, ci, J, r, v0, order, relax)
    assert rc == 0
    assert np.max(np.abs(v - x)) <= 1e-10 * np.max(np.abs(x))


def test_jacobi_order_zero_leaves_the_vector():
    N, pat, J, r = _system()
    v0 = np.random.default_rng(2).standard_normal(3 * N)
    rc, v = orc.jacobi_apply(N, 3, pat["row_ptr"], pat["col_idx"], J, r, v0, 0)
    assert rc == 0 and np.array_equal(v, v0)


def test_jacobi_fixed_point_is_the_solution():
    """Many sweeps on a diagonally dominant system converge to A^-1 d; a
    transposed block or a wrong sign in the off-diagonal sum would not."""
    N, pat, J, r = _system(dominant=True)
    rp, ci = pat["row_ptr"], pat["col_idx"]
    A = _dense(N, rp, ci, J)
    x = np.linalg.solve(A, r)
    rc, v = orc.jacobi_apply(N, 3, rp, ci, J, r, np.zeros(3 * N), 200)
    assert rc == 0
    assert np.max(np.abs(v - x)) <= 1e-12 * np.max(np.abs(x))


def test_bicgstab_with_the_default_jacobi_order_breaks_down():
    """PreconditionerOrder 0 (the reference default): y = K^-1 p leaves y = 0,
    so v = A y = 0 and <r0hat, v> = 0 -- NumericalProblem in the first
    iteration; with three sweeps the solve converges."""
    N, pat, J, r = _system((6, 5, 2), seed=4)
    rp, ci = pat["row_ptr"], pat["col_idx"]
    rc, x, its = orc.bicgstab(N, 3, rp, ci, J, r, orc.params(precond=1, prec_order=0))
    assert rc == 2 and its == 0
    rc, x, its = orc.bicgstab(N, 3, rp, ci, J, r, orc.params(precond=1, prec_order=3, lin_tol=1e-8,
                                                                lin_max_iter=2000))
    assert rc == 0
    A = _dense(N, rp, ci, J)
    assert np.max(np.abs(A @ x - r)) <= 1e-7 * np.max(np.abs(r))
