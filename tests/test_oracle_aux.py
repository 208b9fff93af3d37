"""CPU tests of the oracle's auxiliary equations (BaseAuxiliaryModule,
baseauxiliarymodule.hh:55-121; its pattern entries added by createMatrix_,
tpfalinearizer.hh:503-507; its linearize after the domain's,
linearizeAuxiliaryEquations, :254-281; the Newton loop's treatment of its
DOFs, newtonmethod.hh:287, :583-585, :704-708).  Pinned by: the extended
pattern equal to a brute-force sorted union of the grid's and the module's
entries, the grid's blocks keeping their values in it; the module's terms
added exactly where and in the order given; the extended system's ILU(0)
and BiCGStab against dense algebra; the auxiliary rows left out of the Newton
error and their update current - delta; and a run whose rate-controlled
well (an auxiliary DOF holding the bottom-hole pressure) puts exactly its
target rate of water into the reservoir."""
import numpy as np
import pytest

import oracle as orc
from cases import RateWell, case, mixed_state


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def _well(c, seed=3):
    s = c.spec
    col = [I for I in range(c.N) if I % (s.nx * s.ny) == s.nx + 1]        # cells (1, 1, k)
    wi = np.random.default_rng(seed).uniform(0.5e-9, 2e-9, len(col))
    return RateWell(c.N, [(col, wi, 4.0)])


def _prob(c, well):
    rows, cols = well.links()
    return orc.with_auxiliary(orc.problem_from_case(c), well.n_aux, rows, cols)


def test_pattern_is_the_union_of_grid_and_module_entries():
    c = case(4, 3, 3, seed=1)
    well = RateWell(c.N, [([5, 17, 29], [1e-9] * 3, 1.0), ([0, 35], [2e-9] * 2, 2.0)])
    plain = orc.pattern(orc.problem_from_case(c))
    prob = _prob(c, well)
    pat = orc.pattern(prob)
    R = c.N + 2
    want = [set(plain["col_idx"][plain["row_ptr"][I]:plain["row_ptr"][I + 1]]) for I in range(c.N)]
    want += [set() for _ in range(2)]
    for r, cc in zip(*well.links()):
        want[r].add(int(cc))
    assert len(pat["row_ptr"]) == R + 1
    for r in range(R):
        got = list(pat["col_idx"][pat["row_ptr"][r]:pat["row_ptr"][r + 1]])
        assert got == sorted(want[r])
        assert pat["col_idx"][pat["diag_off"][r]] == r
    # neighbour offsets point at the same blocks A[J][I]
    for I in range(c.N):
        for k in range(pat["nb_ptr"][I], pat["nb_ptr"][I + 1]):
            J = pat["nb_cell"][k]
            assert pat["col_idx"][pat["nb_off"][k]] == I
            assert pat["nb_off"][k] >= pat["row_ptr"][J] and pat["nb_off"][k] < pat["row_ptr"][J + 1]


def test_grid_blocks_keep_their_values_and_module_terms_add_in_order():
    c = case(5, 4, 3, seed=2)
    well = _well(c)
    pv, mean = mixed_state(c, seed=2)
    plain = orc.problem_from_case(c)
    st1 = orc.storage(plain, *mixed_state(c, seed=3))
    J0, r0 = orc.linearize(plain, pv, mean, st1, 43200.0)
    prob = _prob(c, well)
    J1, r1 = orc.linearize(prob, pv, mean, st1, 43200.0)
    p0, p1 = orc.pattern(plain), orc.pattern(prob)
    for I in range(c.N):
        for k in range(p0["row_ptr"][I], p0["row_ptr"][I + 1]):
            col = p0["col_idx"][k]
            kk = p1["row_ptr"][I] + list(p1["col_idx"][p1["row_ptr"][I]:p1["row_ptr"][I + 1]]).index(col)
            assert np.array_equal(_bits(J1[9 * kk:9 * kk + 9]), _bits(J0[9 * k:9 * k + 9]))
    assert np.array_equal(_bits(r1[:3 * c.N]), _bits(r0)) and not np.any(r1[3 * c.N:])
    aux = well.initial(pv)
    brow, bcol, blocks, rrow, rvals = well.terms(pv, mean, aux)
    J2, r2 = orc.add_auxiliary(prob, J1, r1, brow, bcol, blocks, rrow, rvals)
    Jw, rw = J1.copy(), r1.copy()
    for b in range(len(brow)):
        row = p1["col_idx"][p1["row_ptr"][brow[b]]:p1["row_ptr"][brow[b] + 1]]
        k = p1["row_ptr"][brow[b]] + list(row).index(bcol[b])
        Jw[9 * k:9 * k + 9] = Jw[9 * k:9 * k + 9] + blocks[b]
    for b in range(len(rrow)):
        rw[3 * rrow[b]:3 * rrow[b] + 3] = rw[3 * rrow[b]:3 * rrow[b] + 3] + rvals[b]
    assert np.array_equal(_bits(J2), _bits(Jw)) and np.array_equal(_bits(r2), _bits(rw))
    with pytest.raises(AssertionError):           # a block outside the pattern
        orc.add_auxiliary(prob, J1, r1, [c.N], [3], np.ones((1, 9)))


def test_extended_system_solves_like_dense_algebra():
    c = case(3, 3, 2, seed=4)
    well = _well(c)
    pv, mean = mixed_state(c, seed=4)
    prob = _prob(c, well)
    st1 = orc.storage(prob, *mixed_state(c, seed=5))
    J, r = orc.linearize(prob, pv, mean, st1, 43200.0)
    J, r = orc.add_auxiliary(prob, J, r, *well.terms(pv, mean, well.initial(pv)))
    pat = orc.pattern(prob)
    R = orc.rows(prob)
    A = np.zeros((3 * R, 3 * R))
    for i in range(R):
        for k in range(pat["row_ptr"][i], pat["row_ptr"][i + 1]):
            j = pat["col_idx"][k]
            A[3 * i:3 * i + 3, 3 * j:3 * j + 3] = J[9 * k:9 * k + 9].reshape(3, 3)
    x = np.linalg.solve(A, r)
    rc, xo, its = orc.bicgstab(R, 3, pat["row_ptr"], pat["col_idx"], J, r, orc.params(lin_tol=1e-12,
                                                                                       lin_max_iter=500))
    assert rc == 0
    assert np.max(np.abs(xo - x)) <= 1e-8 * np.max(np.abs(x))
    # ILU(0) of the extended system: L U equals A on the pattern
    rc, lu = orc.ilu0(R, 3, pat["row_ptr"], pat["col_idx"], J)
    assert rc == 0
    Lm, Um = np.eye(3 * R), np.zeros((3 * R, 3 * R))
    for i in range(R):
        for k in range(pat["row_ptr"][i], pat["row_ptr"][i + 1]):
            j = pat["col_idx"][k]
            blk = lu[9 * k:9 * k + 9].reshape(3, 3)
            if j < i:
                Lm[3 * i:3 * i + 3, 3 * j:3 * j + 3] = blk
            elif j == i:
                Um[3 * i:3 * i + 3, 3 * j:3 * j + 3] = np.linalg.inv(blk)    # pivots stored inverted
            else:
                Um[3 * i:3 * i + 3, 3 * j:3 * j + 3] = blk
    LU = Lm @ Um
    for i in range(R):
        for k in range(pat["row_ptr"][i], pat["row_ptr"][i + 1]):
            j = pat["col_idx"][k]
            assert np.allclose(LU[3 * i:3 * i + 3, 3 * j:3 * j + 3], A[3 * i:3 * i + 3, 3 * j:3 * j + 3],
                               rtol=1e-10, atol=1e-12 * np.max(np.abs(A)))


def test_run_with_a_rate_well_injects_its_target():
    """Four steps of a 6x5x4 grid without its own wells and with one
    rate-controlled water injector as an auxiliary DOF: converged, the well's
    rate equals its target, the water in place grows by Q times the elapsed
    time and oil and gas stay; the well's pressure is the auxiliary DOF."""
    c = case(6, 5, 4, seed=6)
    well = _well(c, seed=6)
    prob = _prob(c, well)
    prob.n_src = 0
    s = c.spec
    V = s.dx * s.dy * s.dz
    aux0 = well.initial(c.pv0)
    m0 = orc.storage(prob, c.pv0, c.mean0).reshape(-1, 3).sum(axis=0) * V
    iters = []

    def module(pv, mean, aux, it):
        iters.append(it)
        return well.terms(pv, mean, aux, it)

    rc, pv, mean, aux, hist, dts, log = orc.simulate_aux(prob, c.pv0, c.mean0, aux0, 4, 86400.0,
                                                         orc.params(newton_tol=1e-9), module, nthreads=4)
    assert rc == 0
    assert len(iters) == sum(1 for l in log)               # the module after every linearization
    q = sum(WellRate for WellRate in (well.RHO * wi * (aux[0] - pv[3 * k + 1])
                                      for k, wi in zip(*well.wells[0][:2])))
    assert q == pytest.approx(4.0, rel=1e-8)
    m1 = orc.storage(prob, pv, mean).reshape(-1, 3).sum(axis=0) * V
    injected = 4.0 * float(np.sum(dts))
    assert m1[1] - m0[1] == pytest.approx(injected, rel=1e-6)
    assert abs(m1[0] - m0[0]) <= 1e-6 * injected and abs(m1[2] - m0[2]) <= 1e-6 * injected
    assert aux[0] > np.max(pv[3 * well.wells[0][0] + 1])   # injecting: above the cells' pressures
    assert aux[1] == 0.0 and aux[2] == 0.0


def test_failed_module_halves_the_step():
    c = case(4, 4, 2, seed=7)
    well = _well(c, seed=7)
    prob = _prob(c, well)
    calls = []

    def module(pv, mean, aux, it):
        calls.append(it)
        return None if len(calls) == 1 else well.terms(pv, mean, aux, it)

    rc, pv, mean, aux, hist, dts, log = orc.simulate_aux(prob, c.pv0, c.mean0, well.initial(c.pv0), 1, 86400.0,
                                                         orc.params(), module)
    assert rc == 0 and dts[0] == 43200.0
    assert log and all(l[1] == 1 for l in log)             # attempt 0 failed before logging


@pytest.mark.parametrize("what", ["range", "diagonal", "links"])
def test_invalid_auxiliary_patterns_are_rejected(what):
    c = case(3, 2, 2)
    N = c.N
    rows, cols = [N, N, 0], [N, 0, N]
    n = 1
    if what == "range":
        rows = rows + [N + 1]; cols = cols + [0]
    elif what == "diagonal":
        rows, cols = [N, 0], [0, N]
    elif what == "links":
        rows, cols = [], []
    prob = orc.with_auxiliary(orc.problem_from_case(c), n, rows, cols) if rows else orc.problem_from_case(c)
    if not rows:
        prob.n_aux = 1
    with pytest.raises(AssertionError):
        orc.pattern(prob)
