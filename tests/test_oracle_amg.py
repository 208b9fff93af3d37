"""The oracle's aggregation AMG (the product's precond = 2, in place of the
reference's ParallelAmgBackend, parallelamgbackend.hh:85-316), pinned by
dense linear algebra written here independently of the oracle's sweeps:
- a Gauss-Seidel sweep in red-black order (cells with i + j + k even first,
  natural order within a colour) is the block triangular solve of the
  permuted matrix, (D + L) x_new = b - U x_old (forward) and (D + U) x_new =
  b - L x_old (backward), with x_old = 0 for the zero start;
- the coarse operators are P^T A P for the piecewise-constant P of the box
  aggregates;
- the coarsest solve inverts its matrix (numpy's inverse, singular input
  rejected, a zero leading pivot handled by the row exchange);
- a V-cycle over two levels is the composition pre-smoothing, coarse
  correction with the damped prolongation of the exact coarse solution,
  post-smoothing, and over one level it is the exact solve;
- BiCGStab with it converges to the solution of the system, and a Newton run
  with it ends within the linear tolerance's reach of the ILU(0) run."""
import numpy as np
import pytest

import oracle as orc
from cases import case, mixed_state


def red_black_order(nx, ny, nz):
    """cells with i + j + k even, then the odd ones, natural order within each"""
    I = np.arange(nx * ny * nz)
    par = (I % nx + (I // nx) % ny + I // (nx * ny)) % 2
    return np.concatenate([I[par == 0], I[par == 1]])

OFF = lambda nx, ny: (-nx * ny, -nx, -1, 0, 1, nx, nx * ny)


def _system(dims, seed=3):
    c = case(*dims, seed=seed)
    prob = orc.problem_from_case(c)
    pv, mean = mixed_state(c, seed=seed)
    J, r = orc.linearize(prob, pv, mean, orc.storage(prob, *mixed_state(c, seed=seed + 1)), 43200.0)
    return c, orc.pattern(prob), J, r


def _dense_from_dia(dims, A7):
    """dense matrix of a structured-layout [7][N][9] matrix (neighbours by coordinates)"""
    nx, ny, nz = dims
    N = nx * ny * nz
    A7 = A7.reshape(7, N, 3, 3)
    M = np.zeros((3 * N, 3 * N))
    for I in range(N):
        i, j, k = I % nx, (I // nx) % ny, I // (nx * ny)
        nb = [(0, k > 0), (1, j > 0), (2, i > 0), (3, True), (4, i + 1 < nx), (5, j + 1 < ny), (6, k + 1 < nz)]
        for s, ok in nb:
            if ok:
                J = I + OFF(nx, ny)[s]
                M[3 * I:3 * I + 3, 3 * J:3 * J + 3] = A7[s, I]
    return M


def _dense_gs(M, order, b, x_old):
    """block Gauss-Seidel, cells in the given order, as a block triangular
    solve: (D + L) x_new = b - U x_old of the permuted matrix"""
    N = len(order)
    idx = (3 * np.asarray(order)[:, None] + np.arange(3)).reshape(-1)
    P = M[np.ix_(idx, idx)]
    pos = np.kron(np.arange(N)[:, None] - np.arange(N)[None, :], np.ones((3, 3)))
    lower_incl = np.where(pos >= 0, P, 0.0)
    rest = P - lower_incl
    rhs = b[idx] - rest @ x_old[idx]
    x = np.empty(3 * N)
    x[idx] = np.linalg.solve(lower_incl, rhs)
    return x


def _diag_dominant_dia(dims, seed):
    """a random structured-layout matrix whose block rows are diagonally dominant"""
    nx, ny, nz = dims
    N = nx * ny * nz
    rng = np.random.default_rng(seed)
    A7 = rng.uniform(-1.0, 1.0, (7, N, 3, 3))
    A7[3] += np.eye(3) * 25.0
    return A7.reshape(-1)


@pytest.mark.parametrize("dims", [(4, 2, 8), (5, 3, 9), (9, 4, 17), (1, 1, 7), (6, 1, 3)])
@pytest.mark.parametrize("forward,zero", [(True, True), (True, False), (False, False)])
def test_gs_sweep_is_the_block_triangular_solve_in_red_black_order(dims, forward, zero):
    N = int(np.prod(dims))
    A7 = _diag_dominant_dia(dims, seed=N)
    rng = np.random.default_rng(7)
    b = rng.standard_normal(3 * N)
    x0 = np.zeros(3 * N) if zero else rng.standard_normal(3 * N)
    rc, x = orc.gs_sweep(dims, A7, b, x0, forward=forward, zero=zero)
    assert rc == 0
    order = red_black_order(*dims)
    if not forward:
        order = order[::-1]
    want = _dense_gs(_dense_from_dia(dims, A7), order, b, x0)
    assert np.max(np.abs(x - want)) <= 1e-12 * np.max(np.abs(want))


def test_gs_iterations_converge_to_the_solution():
    dims = (6, 5, 9)
    N = int(np.prod(dims))
    A7 = _diag_dominant_dia(dims, seed=3)
    b = np.random.default_rng(3).standard_normal(3 * N)
    x = np.zeros(3 * N)
    for k in range(40):
        _, x = orc.gs_sweep(dims, A7, b, x, forward=k % 2 == 0, zero=k == 0)
    want = np.linalg.solve(_dense_from_dia(dims, A7), b)
    assert np.max(np.abs(x - want)) <= 1e-10 * np.max(np.abs(want))


def _prolongator(fdims, agg):
    nx, ny, nz = fdims
    cx, cy, cz = (-(-nx // agg[0]), -(-ny // agg[1]), -(-nz // agg[2]))
    N, Nc = nx * ny * nz, cx * cy * cz
    P = np.zeros((3 * N, 3 * Nc))
    for I in range(N):
        i, j, k = I % nx, (I // nx) % ny, I // (nx * ny)
        a = i // agg[0] + cx * (j // agg[1] + cy * (k // agg[2]))
        P[3 * I:3 * I + 3, 3 * a:3 * a + 3] = np.eye(3)
    return P, (cx, cy, cz)


@pytest.mark.parametrize("dims,agg", [((9, 6, 12), (3, 3, 3)), ((10, 7, 5), (3, 3, 3)), ((8, 8, 8), (2, 2, 4)),
                                      ((5, 1, 11), (3, 1, 3))])
def test_coarse_operator_is_the_galerkin_product(dims, agg):
    c, pat, J, r = _system(dims)
    prm = orc.amg_params(amg_agg=agg, amg_coarsen_target=1)
    d = orc.amg_dims(dims, prm)
    assert len(d) >= 2
    P, cdims = _prolongator(dims, agg)
    assert tuple(d[1]) == cdims
    rc0, A0 = orc.amg_level_matrix(dims, pat["row_ptr"], pat["col_idx"], J, prm, 0)
    assert np.array_equal(A0, orc.dia_from_bcrs(dims, pat["row_ptr"], pat["col_idx"], J))
    rc1, A1 = orc.amg_level_matrix(dims, pat["row_ptr"], pat["col_idx"], J, prm, 1)
    M = _dense_from_dia(dims, A0)
    want = P.T @ M @ P
    got = _dense_from_dia(cdims, A1)
    scale = np.max(np.abs(want))
    assert np.max(np.abs(got - want)) <= 1e-13 * scale
    # and nothing outside the 7-point pattern of the coarse grid
    assert np.max(np.abs(want - _dense_from_dia(cdims, orc.dia_from_bcrs(cdims, *_pattern_of(cdims), _blocks(
        want, cdims))))) == 0.0


def _pattern_of(dims):
    nx, ny, nz = dims
    N = nx * ny * nz
    rp, ci = [0], []
    for I in range(N):
        i, j, k = I % nx, (I // nx) % ny, I // (nx * ny)
        for s, ok in enumerate([k > 0, j > 0, i > 0, True, i + 1 < nx, j + 1 < ny, k + 1 < nz]):
            if ok:
                ci.append(I + OFF(nx, ny)[s])
        rp.append(len(ci))
    return np.array(rp, np.int32), np.array(ci, np.int32)


def _blocks(M, dims):
    rp, ci = _pattern_of(dims)
    out = []
    for I in range(len(rp) - 1):
        for k in range(rp[I], rp[I + 1]):
            out.append(M[3 * I:3 * I + 3, 3 * ci[k]:3 * ci[k] + 3].reshape(-1))
    return np.concatenate(out)


def test_dense_inverse_matches_numpy_and_pivots():
    rng = np.random.default_rng(11)
    M = rng.standard_normal((45, 45))
    M[0, 0] = 0.0                       # the first pivot needs a row exchange
    rc, Minv = orc.dense_inverse(M)
    assert rc == 0
    want = np.linalg.inv(M)
    assert np.max(np.abs(Minv - want)) <= 1e-11 * np.max(np.abs(want))
    S = M.copy()
    S[7] = 0.0                           # singular (exactly: the row stays zero)
    assert orc.dense_inverse(S)[0] == 2
    Q = M.copy()
    Q[4, 5] = np.nan
    assert orc.dense_inverse(Q)[0] == 2


def test_one_level_cycle_is_the_exact_solve():
    dims = (3, 3, 3)
    c, pat, J, r = _system(dims)
    prm = orc.amg_params(amg_coarsen_target=27)
    assert len(orc.amg_dims(dims, prm)) == 1
    d = np.random.default_rng(5).standard_normal(3 * c.N)
    rc, v = orc.amg_apply(dims, pat["row_ptr"], pat["col_idx"], J, prm, d)
    assert rc == 0
    M = _dense_from_dia(dims, orc.dia_from_bcrs(dims, pat["row_ptr"], pat["col_idx"], J))
    want = np.linalg.solve(M, d)
    assert np.max(np.abs(v - want)) <= 1e-9 * np.max(np.abs(want))


@pytest.mark.parametrize("pre,post,damp", [(2, 2, 1.6), (1, 1, 1.0), (1, 0, 1.3)])
def test_two_level_cycle_is_smoothing_around_the_coarse_correction(pre, post, damp):
    dims, agg = (9, 6, 6), (3, 3, 3)
    c, pat, J, r = _system(dims)
    N = c.N
    prm = orc.amg_params(amg_agg=agg, amg_coarsen_target=24, amg_pre_smooth=pre, amg_post_smooth=post,
                         amg_prolong_damp=damp)
    assert len(orc.amg_dims(dims, prm)) == 2            # 54 fine -> 3x2x2 = 12 coarse cells
    d = np.random.default_rng(9).standard_normal(3 * N)
    rc, v = orc.amg_apply(dims, pat["row_ptr"], pat["col_idx"], J, prm, d)
    assert rc == 0
    M = _dense_from_dia(dims, orc.dia_from_bcrs(dims, pat["row_ptr"], pat["col_idx"], J))
    P, cdims = _prolongator(dims, agg)
    order = red_black_order(*dims)
    x = np.zeros(3 * N)
    for k in range(pre):
        x = _dense_gs(M, order, d, x)
    xc = np.linalg.solve(P.T @ M @ P, P.T @ (d - M @ x))
    x = x + damp * (P @ xc)
    for k in range(post):
        x = _dense_gs(M, order[::-1], d, x)
    assert np.max(np.abs(v - x)) <= 1e-9 * np.max(np.abs(x))


def test_hierarchy_dimensions():
    prm = orc.amg_params()
    d = orc.amg_dims((100, 100, 50), prm)
    assert [tuple(x) for x in d] == [(100, 100, 50), (34, 34, 17), (12, 12, 6), (4, 4, 2)]
    d = orc.amg_dims((1, 1, 5), prm)
    assert [tuple(x) for x in d] == [(1, 1, 5)]
    assert orc.amg_dims((10, 10, 10), orc.amg_params(amg_coarsen_target=33)).size == 0   # beyond the dense limit
    assert orc.amg_dims((10, 10, 10), orc.amg_params(amg_pre_smooth=0)).size == 0


@pytest.mark.parametrize("dims", [(12, 10, 9), (20, 7, 13)])
def test_bicgstab_with_amg_solves_the_system(dims):
    c, pat, J, r = _system(dims, seed=4)
    prm = orc.amg_params(lin_tol=1e-10, lin_abs_tol=0.0)
    rc, x, its = orc.bicgstab_grid(dims, pat["row_ptr"], pat["col_idx"], J, r, prm)
    assert rc == 0
    M = _dense_from_dia(dims, orc.dia_from_bcrs(dims, pat["row_ptr"], pat["col_idx"], J))
    want = np.linalg.solve(M, r)
    assert np.max(np.abs(x - want)) <= 1e-7 * np.max(np.abs(want))
    rc2, x2, its_ilu = orc.bicgstab_grid(dims, pat["row_ptr"], pat["col_idx"], J, r,
                                         orc.params(lin_tol=1e-10, lin_abs_tol=0.0))
    print("BiCGStab iterations: AMG", its, "ILU(0)", its_ilu)
    assert its <= its_ilu


def test_newton_run_with_amg_follows_the_ilu_run():
    c = case(12, 12, 6, seed=21, wells=2)
    prob = orc.problem_from_case(c)
    kw = dict(max_dt=86400.0, min_dt=86400.0, newton_tol=1e-8)
    out = []
    for prm in (orc.params(**kw), orc.amg_params(**kw)):
        rc, pv, mean, hist, dts, log, tim = orc.simulate(prob, c.pv0, c.mean0, 3, 86400.0, prm, nthreads=4)
        assert rc == 0
        out.append((pv, [l[3] for l in log]))
    print("linear iterations per Newton iteration: ILU", out[0][1], "AMG", out[1][1])
    p0, p1 = out[0][0][1::3], out[1][0][1::3]
    assert np.max(np.abs(p1 - p0)) / np.max(np.abs(p0)) <= 1e-7
    assert np.max(np.abs(out[1][0][0::3] - out[0][0][0::3])) <= 1e-6
