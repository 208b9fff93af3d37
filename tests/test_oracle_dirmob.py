"""CPU tests of the oracle's directional mobilities (the face flux takes the
upwind cell's mobility along the face's axis, BlackOilIntensiveQuantities::
mobility(phaseIdx, facedir), blackoilintensivequantities.hh:530-550, used at
blackoillocalresidualtpfa.hh:375-381), with the directional relative
permeabilities kr_dir = m_dir kr.  Pinned by: multipliers of one changing no
bit; a multiplier of two on one axis doubling exactly that axis' face rates
and nothing else; a zero multiplier removing one phase's flow along one
axis; random multipliers scaling each face rate by the upwind cell's value
(the upwind cell read off the sign of the plain rate); central differences
of the Jacobian; invalid input rejected."""
import numpy as np
import pytest

import oracle as orc
from cases import case, mixed_state


COL = (1, 0, 2)      # FLORES column of each phase (water, oil, gas): by component (oil, water, gas)


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def _axis_of_entries(prob):
    """axis (0, 1, 2) of each neighbour entry in the pattern's order"""
    pat = orc.pattern(prob)
    nx, ny = prob.nx, prob.ny
    ax = np.empty(len(pat["nb_cell"]), np.int64)
    for I in range(len(pat["nb_ptr"]) - 1):
        for k in range(pat["nb_ptr"][I], pat["nb_ptr"][I + 1]):
            d = abs(int(pat["nb_cell"][k]) - I)
            ax[k] = 0 if d == 1 and nx > 1 else (1 if d == nx and ny > 1 else 2)
    return pat, ax


def test_unit_multipliers_change_no_bit():
    c = case(5, 4, 3, seed=7)
    pv, mean = mixed_state(c, seed=7)
    plain = orc.problem_from_case(c)
    st1 = orc.storage(plain, *mixed_state(c, seed=8))
    J0, r0 = orc.linearize(plain, pv, mean, st1, 43200.0)
    f0, d0 = orc.flows(plain, pv, mean)
    prob = orc.with_directional_mobility(orc.problem_from_case(c), np.ones(9 * c.N))
    J1, r1 = orc.linearize(prob, pv, mean, st1, 43200.0)
    f1, d1 = orc.flows(prob, pv, mean)
    for a, b in ((J0, J1), (r0, r1), (f0, f1), (d0, d1)):
        assert np.array_equal(_bits(a), _bits(b))


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_doubling_one_axis_doubles_its_rates_exactly(axis):
    c = case(4, 3, 3, seed=11)
    pv, mean = mixed_state(c, seed=11)
    plain = orc.problem_from_case(c)
    pat, ax = _axis_of_entries(plain)
    f0, d0 = orc.flows(plain, pv, mean)
    m = np.ones((3, 3, c.N))
    m[axis] = 2.0
    prob = orc.with_directional_mobility(orc.problem_from_case(c), m)
    f1, d1 = orc.flows(prob, pv, mean)
    on = ax == axis
    assert on.any()
    assert np.array_equal(_bits(d1[on]), _bits(2.0 * d0[on]))
    assert np.array_equal(_bits(f1[on]), _bits(2.0 * f0[on]))
    assert np.array_equal(_bits(d1[~on]), _bits(d0[~on]))
    assert np.array_equal(_bits(f1[~on]), _bits(f0[~on]))


def test_zero_multiplier_stops_one_phase_along_one_axis():
    c = case(4, 4, 3, seed=13)
    pv, mean = mixed_state(c, seed=13)
    plain = orc.problem_from_case(c)
    pat, ax = _axis_of_entries(plain)
    f0, d0 = orc.flows(plain, pv, mean)
    m = np.ones((3, 3, c.N))
    m[2, 0] = 0.0                       # no water along z
    prob = orc.with_directional_mobility(orc.problem_from_case(c), m)
    f1, d1 = orc.flows(prob, pv, mean)
    z = ax == 2
    w = COL[0]
    assert np.all(d1[z, w] == 0.0) and np.any(d0[z, w] != 0.0)
    others = [q for q in range(3) if q != w]
    assert np.array_equal(_bits(d1[z][:, others]), _bits(d0[z][:, others]))
    assert np.array_equal(_bits(d1[~z]), _bits(d0[~z]))


def test_rates_scale_by_the_upwind_cells_multiplier():
    c = case(5, 4, 3, seed=17)
    pv, mean = mixed_state(c, seed=17)
    plain = orc.problem_from_case(c)
    pat, ax = _axis_of_entries(plain)
    f0, d0 = orc.flows(plain, pv, mean)
    m = np.random.default_rng(17).uniform(0.2, 3.0, (3, 3, c.N))
    prob = orc.with_directional_mobility(orc.problem_from_case(c), m)
    f1, d1 = orc.flows(prob, pv, mean)
    checked = 0
    for I in range(c.N):
        for k in range(pat["nb_ptr"][I], pat["nb_ptr"][I + 1]):
            J = int(pat["nb_cell"][k])
            for a in range(3):
                col = COL[a]
                if d0[k, col] == 0.0:
                    continue
                up = I if d0[k, col] > 0.0 else J       # a positive rate leaves I
                assert d1[k, col] == pytest.approx(d0[k, col] * m[ax[k], a, up], rel=1e-14)
                checked += 1
    assert checked > 50


@pytest.mark.parametrize("seed", [1, 2])
def test_jacobian_with_directional_mobility_matches_central_differences(seed):
    c = case(3, 3, 2, seed=seed)
    N = c.N
    pv, mean = mixed_state(c, seed=seed)
    m = np.random.default_rng(seed).uniform(0.1, 4.0, 9 * N)
    prob = orc.with_directional_mobility(orc.problem_from_case(c), m)
    st1 = orc.storage(prob, *mixed_state(c, seed=seed + 10))
    dt = 86400.0
    J, r = orc.linearize(prob, pv, mean, st1, dt)
    pat = orc.pattern(prob)
    steps = [1e-7, 1e-2, 1e-7]
    for col in range(N):
        for var in range(3):
            h = steps[var] if not (var == 2 and mean[col] == 1) else 1e-5
            pp = pv.copy(); pp[3 * col + var] += h
            pm = pv.copy(); pm[3 * col + var] -= h
            _, rp = orc.linearize(prob, pp, mean, st1, dt)
            _, rm = orc.linearize(prob, pm, mean, st1, dt)
            fd = (rp - rm) / (2 * h)
            for row in range(N):
                ks = np.where(pat["col_idx"][pat["row_ptr"][row]:pat["row_ptr"][row + 1]] == col)[0]
                blk = J[9 * (pat["row_ptr"][row] + ks[0]):][:9].reshape(3, 3) if len(ks) else np.zeros((3, 3))
                scale = np.max(np.abs(J[9 * pat["row_ptr"][row]:9 * pat["row_ptr"][row + 1]]).reshape(-1, 3, 3)[:, :, var])
                assert np.max(np.abs(blk[:, var] - fd[3 * row:3 * row + 3])) <= 2e-5 * max(scale, 1e-300) + 1e-9


@pytest.mark.parametrize("bad", [-1.0, np.nan, np.inf])
def test_invalid_multipliers_are_rejected(bad):
    c = case(2, 2, 1)
    m = np.ones(9 * c.N)
    m[5] = bad
    prob = orc.with_directional_mobility(orc.problem_from_case(c), m)
    with pytest.raises(AssertionError):
        orc.flows(prob, c.pv0, c.mean0)
