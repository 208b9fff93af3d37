"""GPU parity of the aggregation AMG (precond = 2, in place of the
reference's ParallelAmgBackend) with the oracle's restatement of it, bit for
bit: the hierarchy's level matrices (Galerkin products), one V-cycle (block
Gauss-Seidel sweeps in red-black order, restriction, damped
prolongation, the coarsest level's dense inverse), BiCGStab with it, and
whole Newton runs, eager and replayed from graphs.  At 500k cells one
V-cycle is compared with the oracle and the solve is checked by its residual.
The FP32-inner form with the AMG (FP32 level matrices and sweeps, FP64
hierarchy and pivots, FP64 outer refinement) is held to the precision form's
bar: pressures within 1e-6 relative of the FP64 oracle per step."""
import numpy as np
import pytest

import omodels_amd as om
import oracle as orc
from cases import case, mixed_state

pytestmark = pytest.mark.gpu

GRIDS = [(9, 6, 12), (13, 7, 5), (1, 1, 20), (5, 1, 11), (40, 40, 10)]


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def _system(dims, seed=3):
    c = case(*dims, seed=seed)
    prob = orc.problem_from_case(c)
    pv, mean = mixed_state(c, seed=seed)
    J, r = orc.linearize(prob, pv, mean, orc.storage(prob, *mixed_state(c, seed=seed + 1)), 43200.0)
    return c, orc.pattern(prob), J, r


def _both(**kw):
    agg = kw.pop("amg_agg", None)
    extra = {} if agg is None else {"amg_agg": agg}
    return om.default_params(precond=2, **kw, **extra), orc.amg_params(**kw, **extra)


@pytest.mark.parametrize("dims", GRIDS)
@pytest.mark.parametrize("kw", [dict(), dict(amg_agg=(2, 2, 4), amg_coarsen_target=8, amg_pre_smooth=1,
                                              amg_post_smooth=3, amg_prolong_damp=1.0)])
def test_hierarchy_and_cycle_bit_exact(dims, kw):
    c, pat, J, r = _system(dims)
    prm, prmo = _both(**kw)
    m = om.Model(c)
    s = om.Solver(m, prm)
    s.prepare(J, r)
    dims_g = s.amg_dims()
    dims_o = orc.amg_dims(dims, prmo)
    print(dims, "levels", [tuple(x) for x in dims_g], flush=True)
    assert np.array_equal(dims_g, dims_o)
    for level in range(len(dims_o)):
        rc, Ao = orc.amg_level_matrix(dims, pat["row_ptr"], pat["col_idx"], J, prmo, level)
        assert rc == 0
        assert np.array_equal(_bits(s.amg_level(level)), _bits(Ao)), level
    d = np.random.default_rng(4).standard_normal(3 * c.N)
    rc, vo = orc.amg_apply(dims, pat["row_ptr"], pat["col_idx"], J, prmo, d)
    assert rc == 0
    for rep in range(2):
        assert np.array_equal(_bits(s.apply_precond(d)), _bits(vo)), rep


@pytest.mark.parametrize("dims", [(13, 7, 5), (40, 40, 10)])
def test_bicgstab_with_amg_bit_exact(dims):
    c, pat, J, r = _system(dims, seed=6)
    m = om.Model(c)
    for kw in (dict(), dict(lin_tol=1e-9, lin_max_iter=3000)):
        prm, prmo = _both(**kw)
        s = om.Solver(m, prm)
        s.prepare(J, r)
        rc, x, its = s.solve()
        rco, xo, itso = orc.bicgstab_grid(dims, pat["row_ptr"], pat["col_idx"], J, r, prmo)
        print(dims, kw, "iterations", its, flush=True)
        assert (rc, its) == (rco, itso) and rc == 0
        assert np.array_equal(_bits(x), _bits(xo))


def test_amg_newton_run_bit_exact_at_reference_settings():
    """40x40x10 for 10 steps at the reference's settings (NewtonTolerance 1e-6,
    adaptive steps) with the AMG: the oracle's step sizes, iteration counts and
    states bit for bit."""
    c = case(40, 40, 10, seed=1234)
    prob = orc.problem_from_case(c)
    prm, prmo = _both()
    m = om.Model(c)
    m.set_solution(c.pv0, c.mean0)
    rc, hist, dts, log, st = m.simulate(prm, 10, 86400.0)
    rco, pvo, mo, histo, dtso, logo, _ = orc.simulate(prob, c.pv0, c.mean0, 10, 86400.0, prmo, nthreads=4)
    print("linear iterations", [l[3] for l in log], flush=True)
    assert rc == 0 and rco == 0
    assert log == logo
    assert np.array_equal(_bits(dts), _bits(dtso))
    assert np.array_equal(_bits(hist), _bits(histo))


def test_amg_graph_replay_is_identical():
    c = case(20, 16, 10, seed=3)
    runs = []
    for graph in (0, 1):
        m = om.Model(c)
        m.set_solution(c.pv0, c.mean0)
        rc, hist, dts, log, st = m.simulate(om.default_params(precond=2, graph=graph), 4, 86400.0)
        assert rc == 0
        runs.append((hist, log))
    assert runs[0][1] == runs[1][1]
    assert np.array_equal(_bits(runs[0][0]), _bits(runs[1][0]))


def test_cycle_and_solve_at_500k():
    """The 100x100x50 black-oil system (four levels down to 4x4x2): one V-cycle
    against the oracle bit for bit, and the BiCGStab solve's true residual
    within the linear tolerance of the right-hand side."""
    dims = (100, 100, 50)
    c = om.BlackOilCase(om.make_spec(*dims, seed=2024))
    prob = orc.problem_from_case(c)
    st1 = orc.storage(prob, c.pv0, c.mean0)
    J, r = orc.linearize(prob, c.pv0, c.mean0, st1, 86400.0, nthreads=16)
    pat = orc.pattern(prob)
    prm, prmo = _both()
    m = om.Model(c)
    s = om.Solver(m, prm)
    s.prepare(J, r)
    assert [tuple(x) for x in s.amg_dims()] == [(100, 100, 50), (34, 34, 17), (12, 12, 6), (4, 4, 2)]
    d = np.random.default_rng(8).standard_normal(3 * c.N)
    rc, vo = orc.amg_apply(dims, pat["row_ptr"], pat["col_idx"], J, prmo, d)
    assert rc == 0
    assert np.array_equal(_bits(s.apply_precond(d)), _bits(vo))
    rc, x, its = s.solve()
    assert rc == 0
    res = r - orc.spmv(c.N, 3, pat["row_ptr"], pat["col_idx"], J, x)
    print("500k: BiCGStab iterations with the AMG", its, "residual reduction",
          np.max(np.abs(res)) / np.max(np.abs(r)), flush=True)
    assert np.max(np.abs(res)) <= 1e-3 * np.max(np.abs(r))


@pytest.mark.parametrize("graph", [0, 1])
def test_fp32_inner_amg_run_matches_the_fp64_oracle(graph):
    """20 one-day steps of the 40x40x10 black-oil workload in the FP32-inner
    form with the AMG: per-step pressures within 1e-6 relative of the FP64
    natural-order oracle, saturations within 1e-6."""
    c = case(40, 40, 10, seed=1234)
    kw = dict(newton_tol=1e-7, max_dt=86400.0, min_dt=86400.0)
    m = om.Model(c)
    m.set_solution(c.pv0, c.mean0)
    rc, hist, dts, log, st = m.simulate(om.default_params(precision=1, precond=2, graph=graph, **kw), 20, 86400.0)
    rco, pvo, mo, histo, dtso, logo, _ = orc.simulate(orc.problem_from_case(c), c.pv0, c.mean0, 20, 86400.0,
                                                      orc.params(**kw), nthreads=4)
    assert rc == 0 and rco == 0
    p, po = hist[:, 1::3], histo[:, 1::3]
    err = np.max(np.abs(p - po), axis=1) / np.max(np.abs(po), axis=1)
    print("FP32-inner AMG: per-step pressure error", err.max(), "linear iterations", st.linear_iters, flush=True)
    assert err.max() <= 1e-6, err
    assert np.max(np.abs(hist[:, 0::3] - histo[:, 0::3])) <= 1e-6


def test_fp32_inner_amg_at_500k_solves_the_system():
    """The 100x100x50 system with the FP32-inner solve and the AMG: the true
    FP64 residual of the refined solution within the linear tolerance."""
    dims = (100, 100, 50)
    c = om.BlackOilCase(om.make_spec(*dims, seed=2024))
    prob = orc.problem_from_case(c)
    st1 = orc.storage(prob, c.pv0, c.mean0)
    J, r = orc.linearize(prob, c.pv0, c.mean0, st1, 86400.0, nthreads=16)
    pat = orc.pattern(prob)
    m = om.Model(c)
    s = om.Solver(m, om.default_params(precond=2, precision=1))
    s.prepare(J, r)
    rc, x, its = s.solve()
    assert rc == 0
    res = r - orc.spmv(c.N, 3, pat["row_ptr"], pat["col_idx"], J, x)
    print("500k FP32-inner AMG: iterations", its, "residual reduction", np.max(np.abs(res)) / np.max(np.abs(r)),
          flush=True)
    assert np.max(np.abs(res)) <= 1e-3 * np.max(np.abs(r))
