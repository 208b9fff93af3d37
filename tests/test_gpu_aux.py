"""GPU parity of auxiliary equations (omodels_model_set_auxiliary and
friends; BaseAuxiliaryModule, baseauxiliarymodule.hh:55-121, in
TpfaLinearizer::createMatrix_ / linearizeAuxiliaryEquations,
tpfalinearizer.hh:254-281, :503-507, and NewtonMethod, newtonmethod.hh:287,
:583-585, :704-708) with the oracle, bit for bit: the extended pattern, the
Jacobian and residual with the module's terms, the ILU(0) factor and the
BiCGStab solution of the extended system, the Newton update of cells and
auxiliary DOFs, and whole runs with the module called inside the Newton loop
(eager and with graph replay); removed again, the plain bits."""
import numpy as np
import pytest

import omodels_amd as om
import oracle as orc
from cases import RateWell, case, mixed_state

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def _wells(c, seed):
    s = c.spec
    rng = np.random.default_rng(seed)
    inj = [I for I in range(c.N) if I % (s.nx * s.ny) == s.nx + 1]
    prod = [I for I in range(c.N) if I % (s.nx * s.ny) == s.nx * s.ny - s.nx - 2]
    return RateWell(c.N, [(inj, rng.uniform(0.5e-9, 2e-9, len(inj)), 4.0),
                          (prod, rng.uniform(0.5e-9, 2e-9, len(prod)), -1.0)])


def _both(c, well):
    rows, cols = well.links()
    m = om.Model(c)
    m.set_auxiliary(well.n_aux, rows, cols)
    prob = orc.with_auxiliary(orc.problem_from_case(c), well.n_aux, rows, cols)
    return m, prob


@pytest.mark.parametrize("dims,seed", [((4, 3, 2), 1), ((12, 9, 5), 2), ((40, 40, 10), 3)])
def test_pattern_system_and_solve_bit_exact(dims, seed):
    c = case(*dims, seed=seed)
    well = _wells(c, seed)
    m, prob = _both(c, well)
    pg, po = m.pattern(), orc.pattern(prob)
    for k in ("row_ptr", "col_idx", "diag_off", "nb_ptr", "nb_cell", "nb_off"):
        assert np.array_equal(pg[k], po[k]), k
    pv, mean = mixed_state(c, seed=seed)
    st1 = orc.storage(prob, *mixed_state(c, seed=seed + 100))
    m.set_solution(pv, mean)
    m.set_storage1(st1)
    aux = well.initial(pv)
    m.set_auxiliary_solution(aux)
    assert np.array_equal(_bits(m.get_auxiliary_solution()), _bits(aux))
    J, r = m.linearize(43200.0, iteration_index=1)
    Jo, ro = orc.linearize(prob, pv, mean, st1, 43200.0, nthreads=16)
    assert np.array_equal(_bits(J), _bits(Jo)) and np.array_equal(_bits(r), _bits(ro))
    terms = well.terms(pv, mean, aux)
    m.add_auxiliary(*terms)
    J, r = m.system()
    Jo, ro = orc.add_auxiliary(prob, Jo, ro, *terms)
    assert np.array_equal(_bits(J), _bits(Jo)) and np.array_equal(_bits(r), _bits(ro))
    R = orc.rows(prob)
    prm = om.default_params()
    s = om.Solver(m, prm)
    s.prepare()
    rc, lu = orc.ilu0(R, 3, po["row_ptr"], po["col_idx"], Jo)
    assert rc == 0 and np.array_equal(_bits(s.ilu()), _bits(lu))
    rc, x, its = s.solve()
    rco, xo, itso = orc.bicgstab(R, 3, po["row_ptr"], po["col_idx"], Jo, ro, orc.params())
    print(dims, "rows", R, "BiCGStab iterations", its, flush=True)
    assert (rc, its) == (rco, itso) and np.array_equal(_bits(x), _bits(xo))
    # Newton update: the cells' by the black-oil rules, the auxiliary DOFs current - delta
    urc, nsw, sw = m.newton_update(x, prm)
    pvg, mg = m.get_solution()
    rco, nxt, nm, ws, nso = orc.update_primary(prob, pv, mean, xo[:3 * c.N], np.zeros(c.N, np.uint8), orc.params())
    assert urc == 0 and np.array_equal(_bits(pvg), _bits(nxt)) and np.array_equal(mg, nm)
    assert np.array_equal(_bits(m.get_auxiliary_solution()), _bits(aux - xo[3 * c.N:]))
    # removed again: the grid's pattern and bits
    m.set_auxiliary(0)
    m.set_solution(pv, mean)
    J0, r0 = m.linearize(43200.0, iteration_index=1)
    Jp, rp = orc.linearize(orc.problem_from_case(c), pv, mean, st1, 43200.0, nthreads=16)
    assert np.array_equal(_bits(J0), _bits(Jp)) and np.array_equal(_bits(r0), _bits(rp))


@pytest.mark.parametrize("graph", [0, 1])
def test_run_with_rate_wells_bit_exact(graph):
    """Five one-day steps of a 20x16x8 case whose wells are two
    rate-controlled auxiliary DOFs (a water injector and a producer), the
    module called after every linearization of the Newton loop: the oracle's
    step sizes, iteration counts, states and well pressures bit for bit."""
    c = case(20, 16, 8, seed=9)
    well = _wells(c, 9)
    m, prob = _both(c, well)
    prob.n_src = 0
    c.prob.n_src = 0
    m = om.Model(c)
    rows, cols = well.links()
    m.set_auxiliary(well.n_aux, rows, cols)
    m.set_solution(c.pv0, c.mean0)
    aux0 = well.initial(c.pv0)
    m.set_auxiliary_solution(aux0)
    calls = []

    def module(model, it):
        pv, mean = model.get_solution()
        calls.append(it)
        model.add_auxiliary(*well.terms(pv, mean, model.get_auxiliary_solution(), it))
        return 0

    m.set_auxiliary_module(module)
    rc, hist, dts, log, st = m.simulate(om.default_params(graph=graph), 5, 86400.0)
    rco, pvo, mo, auxo, histo, dtso, logo = orc.simulate_aux(
        prob, c.pv0, c.mean0, aux0, 5, 86400.0, orc.params(), lambda pv, mean, aux, it: well.terms(pv, mean, aux, it),
        nthreads=8)
    print("Newton iterations", st.newton_iters, "linear", st.linear_iters, "module calls", len(calls), flush=True)
    assert rc == 0 and rco == 0
    assert log == logo and len(calls) == len(logo)
    assert np.array_equal(_bits(dts), _bits(dtso))
    assert np.array_equal(_bits(hist), _bits(histo))
    assert np.array_equal(_bits(m.get_auxiliary_solution()), _bits(auxo))


def test_failing_module_halves_the_step_like_the_oracle():
    c = case(8, 8, 4, seed=11)
    well = _wells(c, 11)
    m, prob = _both(c, well)
    m.set_solution(c.pv0, c.mean0)
    aux0 = well.initial(c.pv0)
    m.set_auxiliary_solution(aux0)
    n = {"g": 0, "o": 0}

    def module(model, it):
        n["g"] += 1
        if n["g"] == 2:
            return 1
        pv, mean = model.get_solution()
        model.add_auxiliary(*well.terms(pv, mean, model.get_auxiliary_solution(), it))
        return 0

    def module_o(pv, mean, aux, it):
        n["o"] += 1
        return None if n["o"] == 2 else well.terms(pv, mean, aux, it)

    m.set_auxiliary_module(module)
    rc, hist, dts, log, st = m.simulate(om.default_params(), 2, 86400.0)
    rco, pvo, mo, auxo, histo, dtso, logo = orc.simulate_aux(prob, c.pv0, c.mean0, aux0, 2, 86400.0, orc.params(),
                                                             module_o)
    assert rc == 0 and rco == 0 and st.failed_attempts == 1
    assert log == logo
    assert np.array_equal(_bits(dts), _bits(dtso)) and np.array_equal(_bits(hist), _bits(histo))
    assert np.array_equal(_bits(m.get_auxiliary_solution()), _bits(auxo))


@pytest.mark.parametrize("what", ["range", "diagonal", "outside", "no_module", "cleared_module", "amg",
                                  "stale_solver"])
def test_invalid_use_is_rejected(what):
    c = case(3, 2, 2)
    N = c.N
    m = om.Model(c)
    if what == "range":
        with pytest.raises(om.OmodelsError):
            m.set_auxiliary(1, [N, N + 1], [N, 0])
    elif what == "diagonal":
        with pytest.raises(om.OmodelsError):
            m.set_auxiliary(1, [N, 0], [0, N])
    elif what == "outside":
        m.set_auxiliary(1, [N, N, 0], [N, 0, N])
        m.set_solution(c.pv0, c.mean0)
        m.linearize(86400.0)
        with pytest.raises(om.OmodelsError):
            m.add_auxiliary([N], [3], np.ones((1, 9)))
    elif what == "no_module":
        m.set_auxiliary(1, [N, N, 0], [N, 0, N])
        m.set_solution(c.pv0, c.mean0)
        with pytest.raises(om.OmodelsError):
            m.simulate(om.default_params(), 1, 86400.0)
    elif what == "cleared_module":
        m.set_auxiliary(1, [N, N, 0], [N, 0, N])
        m.set_solution(c.pv0, c.mean0)
        m.set_auxiliary_module(lambda model, it: 0)
        m.set_auxiliary_module(None)
        with pytest.raises(om.OmodelsError):
            m.simulate(om.default_params(), 1, 86400.0)
    elif what == "amg":
        m.set_auxiliary(1, [N, N, 0], [N, 0, N])
        with pytest.raises(om.OmodelsError):
            om.Solver(m, om.default_params(precond=2))
    elif what == "stale_solver":       # created before the system grew
        s = om.Solver(m, om.default_params())
        m.set_auxiliary(1, [N, N, 0], [N, 0, N])
        m.set_solution(c.pv0, c.mean0)
        m.linearize(86400.0)
        with pytest.raises(om.OmodelsError):
            s.prepare()
