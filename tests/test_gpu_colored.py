"""GPU parity of the two-colour tile ILU(0) (ilu_order 1) with the oracle's
restatement of it: preconditioner, BiCGStab and whole Newton runs bit for bit
in FP64; the FP32-inner form within the stated tolerance of the FP64
natural-order oracle."""
import numpy as np
import pytest

import omodels_amd as om
import oracle as orc
from cases import case, mixed_state

pytestmark = pytest.mark.gpu

GRIDS = [(4, 2, 8), (9, 5, 11), (13, 3, 17), (1, 1, 20), (6, 7, 3), (40, 40, 10)]


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def _system(dims, seed=3):
    c = case(*dims, seed=seed)
    prob = orc.problem_from_case(c)
    pv, mean = mixed_state(c, seed=seed)
    J, r = orc.linearize(prob, pv, mean, orc.storage(prob, *mixed_state(c, seed=seed + 1)), 43200.0)
    return c, orc.pattern(prob), J, r


@pytest.mark.parametrize("relax", [1.0, 0.9])
@pytest.mark.parametrize("dims", GRIDS)
def test_coloured_preconditioner_bit_exact(dims, relax):
    c, pat, J, r = _system(dims)
    m = om.Model(c)
    s = om.Solver(m, om.default_params(ilu_order=1, prec_relax=relax))
    s.prepare(J, r)
    d = np.random.default_rng(4).standard_normal(3 * c.N)
    for rep in range(2):
        v = s.apply_precond(d)
        rc, vo = orc.colored_ilu_apply(dims, pat["row_ptr"], pat["col_idx"], J, d, relax)
        assert rc == 0
        assert np.array_equal(_bits(v), _bits(vo)), rep


@pytest.mark.parametrize("dims", [(9, 5, 11), (40, 40, 10)])
def test_coloured_bicgstab_bit_exact(dims):
    c, pat, J, r = _system(dims, seed=6)
    m = om.Model(c)
    for kw in (dict(), dict(lin_tol=1e-9, lin_max_iter=3000)):
        s = om.Solver(m, om.default_params(ilu_order=1, **kw))
        s.prepare(J, r)
        rc, x, its = s.solve()
        rco, xo, itso = orc.bicgstab_grid(dims, pat["row_ptr"], pat["col_idx"], J, r, orc.params(ilu_order=1, **kw))
        assert (rc, its) == (rco, itso) and rc == 0
        assert np.array_equal(_bits(x), _bits(xo))


def test_coloured_newton_run_bit_exact_at_reference_settings():
    """Config 2 (40x40x10) for 10 steps at the reference's settings with the
    two-colour ILU(0): the oracle's step sizes, iteration counts and states
    bit for bit; and within 1e-6 of the natural-order run."""
    c = case(40, 40, 10, seed=1234)
    prob = orc.problem_from_case(c)
    m = om.Model(c)
    m.set_solution(c.pv0, c.mean0)
    rc, hist, dts, log, st = m.simulate(om.default_params(ilu_order=1), 10, 86400.0)
    rco, pvo, mo, histo, dtso, logo, _ = orc.simulate(prob, c.pv0, c.mean0, 10, 86400.0, orc.params(ilu_order=1),
                                                      nthreads=4)
    assert rc == 0 and rco == 0
    assert log == logo
    assert np.array_equal(_bits(dts), _bits(dtso))
    assert np.array_equal(_bits(hist), _bits(histo))


def test_coloured_fp32_inner_run_matches_the_fp64_oracle():
    """The FP32-inner form with the two-colour ILU(0) (FP32 off-diagonal
    blocks, FP64 pivots, FP64 outer refinement), 20 one-day steps of config 2:
    per-step pressures within 1e-6 relative of the FP64 natural-order oracle
    (the tolerance of the precision form), saturations within 1e-6."""
    c = case(40, 40, 10, seed=1234)
    kw = dict(newton_tol=1e-7, max_dt=86400.0, min_dt=86400.0)
    m = om.Model(c)
    m.set_solution(c.pv0, c.mean0)
    rc, hist, dts, log, st = m.simulate(om.default_params(precision=1, ilu_order=1, **kw), 20, 86400.0)
    rco, pvo, mo, histo, dtso, logo, _ = orc.simulate(orc.problem_from_case(c), c.pv0, c.mean0, 20, 86400.0,
                                                      orc.params(**kw), nthreads=4)
    assert rc == 0 and rco == 0
    p, po = hist[:, 1::3], histo[:, 1::3]
    err = np.max(np.abs(p - po), axis=1) / np.max(np.abs(po), axis=1)
    assert err.max() <= 1e-6, err
    assert np.max(np.abs(hist[:, 0::3] - histo[:, 0::3])) <= 1e-6


@pytest.mark.parametrize("precision", [0, 1])
def test_coloured_graph_replay_is_identical(precision):
    c = case(20, 16, 10, seed=3)
    runs = []
    for graph in (0, 1):
        m = om.Model(c)
        m.set_solution(c.pv0, c.mean0)
        rc, hist, dts, log, st = m.simulate(om.default_params(ilu_order=1, precision=precision, graph=graph), 4,
                                            86400.0)
        assert rc == 0
        runs.append((hist, log))
    assert runs[0][1] == runs[1][1]
    assert np.array_equal(_bits(runs[0][0]), _bits(runs[1][0]))
