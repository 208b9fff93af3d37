"""Mixed precision form: FP32 Jacobian copy, FP32 ILU(0) factors with FP64
inverted pivots and an FP32 inner BiCGStab, inside FP64 iterative refinement
with FP64 residual, Newton update and norms.  Bar: per-step pressures within
1e-6 relative of the FP64 oracle (and of the FP64 GPU run)."""
import numpy as np
import pytest

import omodels_amd as om
import oracle as orc
from cases import case, mixed_state, rel_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dims", [(6, 5, 1), (12, 10, 4), (40, 40, 10)])
def test_mixed_precision_linear_solve_meets_the_fp64_criterion(dims):
    c = case(*dims, seed=4)
    pv, mean = mixed_state(c, seed=4)
    prob = orc.problem_from_case(c)
    st1 = orc.storage(prob, *mixed_state(c, seed=54))
    J, r = orc.linearize(prob, pv, mean, st1, 86400.0)
    pat = orc.pattern(prob)
    m = om.Model(c)
    s = om.Solver(m, om.default_params(precision=1))
    s.prepare(J, r)
    rc, x, its = s.solve()
    assert rc == 0
    res = r - orc.spmv(c.N, 3, pat["row_ptr"], pat["col_idx"], J, x)
    # the FP64 outer test of CombinedCriterion: reduction by LinearSolverTolerance
    assert np.abs(res).max() <= 1e-3 * np.abs(r).max()
    # FP64 refinement recovers FP64 accuracy from FP32 inner solves
    s2 = om.Solver(m, om.default_params(precision=1, lin_tol=1e-10, lin_abs_tol=0.0))
    s2.prepare(J, r)
    rc2, x2, its2 = s2.solve()
    assert rc2 == 0
    res2 = r - orc.spmv(c.N, 3, pat["row_ptr"], pat["col_idx"], J, x2)
    assert np.abs(res2).max() <= 1e-10 * np.abs(r).max()
    rco, xo, itso = orc.bicgstab(c.N, 3, pat["row_ptr"], pat["col_idx"], J, r,
                                 orc.params(lin_tol=1e-12, lin_abs_tol=0.0))
    assert rel_err(x2, xo) <= 1e-6


def test_mixed_precision_run_matches_fp64_oracle_config2():
    """20 one-day steps of the 40x40x10 black-oil case.  Both runs converge to
    NewtonTolerance 1e-7 kg/s; at the reference's 1e-6 the converged states of
    any two solvers already differ by up to ~1e-6 relative in pressure (the
    width of the Newton tolerance band), so equal accuracy is asserted one
    decade tighter."""
    c = case(40, 40, 10, seed=1234)
    kw = dict(max_dt=86400.0, min_dt=86400.0, newton_tol=1e-7)
    m = om.Model(c)
    m.set_solution(c.pv0, c.mean0)
    rc, hist, dts, log, st = m.simulate(om.default_params(precision=1, **kw), 20, 86400.0)
    assert rc == 0
    rco, pvo, mo, histo, dtso, logo, tim = orc.simulate(orc.problem_from_case(c), c.pv0, c.mean0, 20, 86400.0,
                                                        orc.params(**kw), nthreads=4)
    assert rco == 0
    p, po = hist[:, 1::3], histo[:, 1::3]
    err = np.max(np.abs(p - po), axis=1) / np.max(np.abs(po), axis=1)
    assert err.max() <= 1e-6, err
    assert np.max(np.abs(hist[:, 0::3] - histo[:, 0::3])) <= 1e-5
