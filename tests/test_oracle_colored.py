"""The oracle's block ILU(0) in the two-colour tile ordering (the product's
ilu_order 1): tiles of 4x2x8 cells coloured (tx + ty + tz) mod 2, colour 0
first, natural order inside a tile.  Pinned by the textbook ILU(0) of the
permuted matrix (the oracle's natural-order restatement applied to P A P^T,
which stores L = A D^-1 and keeps every fill-free update), by the natural
order on a grid of one tile, and by convergence of BiCGStab and Newton runs
to the natural-order results."""
import numpy as np
import pytest

import oracle as orc
from cases import case, mixed_state

TX, TY, TZ = 4, 2, 8


def coloured_order(nx, ny, nz):
    """cells in the ordering: colour-0 tiles, then colour-1, natural inside a tile"""
    N = nx * ny * nz
    I = np.arange(N)
    i, j, k = I % nx, (I // nx) % ny, I // (nx * ny)
    ti, tj, tk = i // TX, j // TY, k // TZ
    ntx, nty = -(-nx // TX), -(-ny // TY)
    tile = ti + ntx * (tj + nty * tk)
    col = (ti + tj + tk) % 2
    return np.lexsort((I, tile, col))


def _system(dims, seed=3):
    c = case(*dims, seed=seed)
    prob = orc.problem_from_case(c)
    pv, mean = mixed_state(c, seed=seed)
    J, r = orc.linearize(prob, pv, mean, orc.storage(prob, *mixed_state(c, seed=seed + 1)), 43200.0)
    return c, orc.pattern(prob), J, r


def _permute(N, rp, ci, J, order):
    new_of_old = np.empty(N, np.int64)
    new_of_old[order] = np.arange(N)
    B = J.reshape(-1, 3, 3)
    nrp, nci, nJ = [0], [], []
    for old in order:
        cols = ci[rp[old]:rp[old + 1]]
        nc = new_of_old[cols]
        o = np.argsort(nc)
        nci.extend(nc[o])
        nJ.append(B[rp[old]:rp[old + 1]][o])
        nrp.append(len(nci))
    return (np.array(nrp, np.int32), np.array(nci, np.int32), np.concatenate(nJ).reshape(-1), new_of_old)


@pytest.mark.parametrize("dims", [(4, 2, 8), (9, 5, 11), (13, 3, 17), (1, 1, 20), (8, 4, 16), (6, 7, 3)])
def test_coloured_ilu_is_ilu0_of_the_permuted_matrix(dims):
    c, pat, J, r = _system(dims)
    N = c.N
    order = coloured_order(*dims)
    prp, pci, pJ, new_of_old = _permute(N, pat["row_ptr"], pat["col_idx"], J, order)
    rc, lu = orc.ilu0(N, 3, prp, pci, pJ)
    assert rc == 0
    d = np.random.default_rng(1).standard_normal(3 * N)
    dp = d.reshape(-1, 3)[order].reshape(-1)
    vp = orc.ilu0_apply(N, 3, prp, pci, lu, dp)
    want = np.empty_like(vp).reshape(-1, 3)
    want[order] = vp.reshape(-1, 3)
    rc, got = orc.colored_ilu_apply(dims, pat["row_ptr"], pat["col_idx"], J, d)
    assert rc == 0
    assert np.max(np.abs(got - want.reshape(-1))) <= 1e-10 * np.max(np.abs(want))


def test_coloured_ilu_on_one_tile_is_the_natural_order():
    c, pat, J, r = _system((4, 2, 8))
    rc, lu = orc.ilu0(c.N, 3, pat["row_ptr"], pat["col_idx"], J)
    d = np.random.default_rng(2).standard_normal(3 * c.N)
    v = orc.ilu0_apply(c.N, 3, pat["row_ptr"], pat["col_idx"], lu, d)
    rc, got = orc.colored_ilu_apply((4, 2, 8), pat["row_ptr"], pat["col_idx"], J, d)
    assert np.max(np.abs(got - v)) <= 1e-12 * np.max(np.abs(v))


def test_coloured_order_alternates_between_face_neighbour_tiles():
    nx, ny, nz = 12, 6, 24
    order = coloured_order(nx, ny, nz)
    pos = np.empty(nx * ny * nz, np.int64)
    pos[order] = np.arange(order.size)
    I = np.arange(nx * ny * nz)
    i, j, k = I % nx, (I // nx) % ny, I // (nx * ny)
    col = (i // TX + j // TY + k // TZ) % 2
    assert np.all(pos[col == 0] < np.min(pos[col == 1]))
    for step, n, coord in ((1, nx, i), (nx, ny, j), (nx * ny, nz, k)):
        m = coord + 1 < n
        same_tile = {1: i // TX, nx: j // TY, nx * ny: k // TZ}[step]
        nb_tile = {1: (i + 1) // TX, nx: (j + 1) // TY, nx * ny: (k + 1) // TZ}[step]
        cross = m & (same_tile != nb_tile)
        assert np.all(col[I[cross]] != col[I[cross] + step])


def test_coloured_bicgstab_converges_like_the_natural_order():
    c, pat, J, r = _system((20, 16, 10), seed=5)
    kw = dict(lin_tol=1e-8, lin_max_iter=2000)
    rc0, x0, it0 = orc.bicgstab(c.N, 3, pat["row_ptr"], pat["col_idx"], J, r, orc.params(**kw))
    rc1, x1, it1 = orc.bicgstab_grid((20, 16, 10), pat["row_ptr"], pat["col_idx"], J, r,
                                     orc.params(ilu_order=1, **kw))
    assert rc0 == 0 and rc1 == 0
    assert it1 <= 2 * it0 + 5
    res = r - orc.spmv(c.N, 3, pat["row_ptr"], pat["col_idx"], J, x1)
    assert np.abs(res).max() <= 1e-8 * np.abs(r).max()
    assert np.max(np.abs(x1 - x0)) <= 1e-6 * np.max(np.abs(x0))


def test_coloured_newton_run_reaches_the_natural_order_states():
    c = case(12, 10, 6, seed=8)
    prob = orc.problem_from_case(c)
    kw = dict(newton_tol=1e-10, max_dt=86400.0, min_dt=86400.0)
    rc0, pv0, m0, h0, dt0, log0, _ = orc.simulate(prob, c.pv0, c.mean0, 4, 86400.0, orc.params(**kw))
    rc1, pv1, m1, h1, dt1, log1, _ = orc.simulate(prob, c.pv0, c.mean0, 4, 86400.0, orc.params(ilu_order=1, **kw))
    assert rc0 == 0 and rc1 == 0
    p0, p1 = h0[:, 1::3], h1[:, 1::3]
    assert np.max(np.abs(p1 - p0)) <= 1e-8 * np.max(np.abs(p0))
    assert np.max(np.abs(h1[:, 0::3] - h0[:, 0::3])) <= 1e-8
