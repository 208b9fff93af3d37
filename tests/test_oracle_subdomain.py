"""CPU tests of the oracle's sub-domain linearization (linearizeDomain(domain),
tpfalinearizer.hh:226-245: resetSystem_(domain) at :370-379, linearize_ over
the domain's cells at :711-842, the boundary loop over all faces after it).
Pinned against the full linearization: all cells give its bits; on a part,
the domain's residual rows, diagonal blocks and their blocks towards domain
cells are its bits, their blocks towards cells outside are zero; rows outside
keep their residual and their blocks but for the domain cells' column
blocks, which are the previous value plus the full linearization's block;
invalid domains are rejected."""
import numpy as np
import pytest

import oracle as orc
from cases import case, mixed_state


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def _setup(seed=3, dims=(5, 4, 3)):
    c = case(*dims, seed=seed, wells=1)
    prob = orc.problem_from_case(c)
    pvA, mA = mixed_state(c, seed=seed)
    pvB, mB = mixed_state(c, seed=seed + 1)
    st1 = orc.storage(prob, *mixed_state(c, seed=seed + 2))
    JA, rA = orc.linearize(prob, pvA, mA, st1, 43200.0)       # the previous linearization
    JB, rB = orc.linearize(prob, pvB, mB, st1, 43200.0)       # the full one at the new state
    return c, prob, pvB, mB, st1, JA, rA, JB, rB


def test_all_cells_give_the_full_linearization():
    c, prob, pv, mean, st1, JA, rA, JB, rB = _setup()
    J, r = orc.linearize_domain(prob, pv, mean, st1, 43200.0, np.random.default_rng(1).permutation(c.N), JA, rA)
    assert np.array_equal(_bits(J), _bits(JB)) and np.array_equal(_bits(r), _bits(rB))


@pytest.mark.parametrize("seed", [3, 4])
def test_part_of_the_grid(seed):
    c, prob, pv, mean, st1, JA, rA, JB, rB = _setup(seed)
    rng = np.random.default_rng(seed)
    cells = np.sort(rng.choice(c.N, c.N // 3, replace=False))
    dom = np.zeros(c.N, bool)
    dom[cells] = True
    J, r = orc.linearize_domain(prob, pv, mean, st1, 43200.0, cells, JA, rA)
    pat = orc.pattern(prob)
    rp, ci = pat["row_ptr"], pat["col_idx"]
    for I in range(c.N):
        if dom[I]:
            assert np.array_equal(_bits(r[3 * I:3 * I + 3]), _bits(rB[3 * I:3 * I + 3]))
        else:
            assert np.array_equal(_bits(r[3 * I:3 * I + 3]), _bits(rA[3 * I:3 * I + 3]))
        for k in range(rp[I], rp[I + 1]):
            col = ci[k]
            got = J[9 * k:9 * k + 9]
            if dom[I]:
                want = JB[9 * k:9 * k + 9] if dom[col] else np.zeros(9)
            elif dom[col]:
                want = JA[9 * k:9 * k + 9] + JB[9 * k:9 * k + 9]      # the column block added to the old row
            else:
                want = JA[9 * k:9 * k + 9]
            assert np.array_equal(_bits(got), _bits(want)), (I, col)


@pytest.mark.parametrize("cells", [[], [0, 0], [-1], [10 ** 6]])
def test_invalid_domains_are_rejected(cells):
    c, prob, pv, mean, st1, JA, rA, JB, rB = _setup()
    with pytest.raises(AssertionError):
        orc.linearize_domain(prob, pv, mean, st1, 43200.0, cells, JA, rA)
