"""GPU parity of the model's optional terms together: boundary conditions,
rock compressibility, directional mobilities and rate-controlled wells as
auxiliary DOFs in one Newton run (bit for bit with the oracle in FP64; in the
FP32-inner form, pressures within 1e-6 relative of the FP64 oracle per
step), and a sub-domain linearization of a system with auxiliary rows."""
import numpy as np
import pytest

import omodels_amd as om
import oracle as orc
from cases import RateWell, case, mixed_state

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def _setup(c, seed):
    s = c.spec
    rng = np.random.default_rng(seed)
    comp = rng.uniform(1e-10, 1e-9, c.N)
    mult = np.ones((3, 3, c.N))
    mult[2] = 0.5
    west = [I for I in range(c.N) if I % s.nx == 0]
    state = np.stack([c.pv0[3 * I:3 * I + 3] for I in west])
    state[:, 1] += 2e5
    bc = dict(cell=west, face=[0] * len(west), type=[2] * len(west), state=state, meaning=c.mean0[west])
    inj = [I for I in range(c.N) if I % (s.nx * s.ny) == s.nx * (s.ny // 2) + s.nx // 2]
    well = RateWell(c.N, [(inj, rng.uniform(0.5e-9, 2e-9, len(inj)), 3.0)])
    rows, cols = well.links()
    m = om.Model(c)
    m.set_rock(comp, None, None)
    m.set_directional_mobility(mult)
    m.set_boundary(**bc)
    m.set_auxiliary(well.n_aux, rows, cols)
    prob = orc.with_rock(orc.problem_from_case(c), comp, None, None)
    prob = orc.with_directional_mobility(prob, mult)
    prob = orc.with_boundary(prob, **bc)
    prob = orc.with_auxiliary(prob, well.n_aux, rows, cols)
    return m, prob, well


def _module(well):
    def module(model, it):
        pv, mean = model.get_solution()
        model.add_auxiliary(*well.terms(pv, mean, model.get_auxiliary_solution(), it))
        return 0
    return module


@pytest.mark.parametrize("precision", [0, 1])
def test_newton_run_with_every_optional_term(precision):
    c = case(16, 12, 6, seed=31)
    m, prob, well = _setup(c, 31)
    m.set_solution(c.pv0, c.mean0)
    aux0 = well.initial(c.pv0)
    m.set_auxiliary_solution(aux0)
    m.set_auxiliary_module(_module(well))
    kw = dict(newton_tol=1e-7) if precision else {}
    rc, hist, dts, log, st = m.simulate(om.default_params(precision=precision, **kw), 4, 86400.0)
    rco, pvo, mo, auxo, histo, dtso, logo = orc.simulate_aux(
        prob, c.pv0, c.mean0, aux0, 4, 86400.0, orc.params(**kw), lambda pv, mean, aux, it: well.terms(pv, mean, aux, it),
        nthreads=8)
    print("precision", precision, "Newton iterations", st.newton_iters, "linear", st.linear_iters, flush=True)
    assert rc == 0 and rco == 0
    if precision == 0:
        assert log == logo
        assert np.array_equal(_bits(dts), _bits(dtso)) and np.array_equal(_bits(hist), _bits(histo))
        assert np.array_equal(_bits(m.get_auxiliary_solution()), _bits(auxo))
    else:
        assert np.array_equal(dts, dtso)
        p, po = hist[:, 1::3], histo[:, 1::3]
        err = np.max(np.abs(p - po), axis=1) / np.max(np.abs(po), axis=1)
        print("FP32-inner per-step pressure error", err, flush=True)
        assert err.max() <= 1e-6


def test_subdomain_of_a_system_with_auxiliary_rows():
    c = case(10, 8, 4, seed=32)
    m, prob, well = _setup(c, 32)
    pv, mean = mixed_state(c, seed=32)
    st1 = orc.storage(prob, *mixed_state(c, seed=33))
    m.set_solution(pv, mean)
    m.set_storage1(st1)
    J, r = m.linearize(43200.0, iteration_index=1)
    Jo, ro = orc.linearize(prob, pv, mean, st1, 43200.0, nthreads=8)
    assert np.array_equal(_bits(J), _bits(Jo)) and np.array_equal(_bits(r), _bits(ro))
    pv2, mean2 = mixed_state(c, seed=34)
    m.set_solution(pv2, mean2)
    cells = np.random.default_rng(32).choice(c.N, c.N // 2, replace=False)
    J, r = m.linearize_subdomain(43200.0, cells)
    Jo, ro = orc.linearize_domain(prob, pv2, mean2, st1, 43200.0, cells, Jo, ro)
    assert np.array_equal(_bits(J), _bits(Jo)) and np.array_equal(_bits(r), _bits(ro))


def test_amg_run_with_boundary_rock_and_directional_mobility():
    """The AMG backend (precond = 2) over a model with boundary conditions,
    rock terms and directional mobilities: four steps bit for bit with the
    oracle's AMG run."""
    c = case(18, 15, 9, seed=35, wells=2)
    s = c.spec
    rng = np.random.default_rng(35)
    comp = rng.uniform(1e-10, 1e-9, c.N)
    mult = rng.uniform(0.3, 2.0, (3, 3, c.N))
    top = [I for I in range(c.N) if I < s.nx * s.ny]
    state = np.stack([c.pv0[3 * I:3 * I + 3] for I in top])
    state[:, 1] -= 1e5
    bc = dict(cell=top, face=[4] * len(top), type=[3] * len(top), state=state, meaning=c.mean0[top])
    m = om.Model(c)
    m.set_rock(comp, None, None)
    m.set_directional_mobility(mult)
    m.set_boundary(**bc)
    prob = orc.with_boundary(orc.with_directional_mobility(orc.with_rock(orc.problem_from_case(c), comp, None, None),
                                                           mult), **bc)
    m.set_solution(c.pv0, c.mean0)
    rc, hist, dts, log, st = m.simulate(om.default_params(precond=2), 4, 86400.0)
    rco, pvo, mo, histo, dtso, logo, _ = orc.simulate(prob, c.pv0, c.mean0, 4, 86400.0, orc.amg_params(), nthreads=8)
    print("AMG: Newton iterations", st.newton_iters, "linear", st.linear_iters, flush=True)
    assert rc == 0 and rco == 0
    assert log == logo
    assert np.array_equal(_bits(dts), _bits(dtso)) and np.array_equal(_bits(hist), _bits(histo))
