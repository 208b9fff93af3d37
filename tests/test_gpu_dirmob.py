"""GPU parity of directional mobilities (omodels_model_set_directional_mobility;
BlackOilIntensiveQuantities::mobility(phaseIdx, facedir),
blackoilintensivequantities.hh:530-550, in the face flux at
blackoillocalresidualtpfa.hh:375-381): residual, Jacobian, FLOWS/FLORES and
simulated states bit-exact with the oracle, alone and together with rock
compressibility, a caller-given discretization and the diffusive term; off
again, the plain bits."""
import numpy as np
import pytest

import omodels_amd as om
import oracle as orc
from cases import case, mixed_state, random_discretization

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def _mult(c, seed):
    m = np.random.default_rng(seed).uniform(0.05, 4.0, (3, 3, c.N))
    m[2, 2, ::7] = 0.0          # no gas along z in some cells
    return m


@pytest.mark.parametrize("dims,seed", [((1, 1, 1), 1), ((4, 3, 2), 2), ((33, 9, 3), 3), ((100, 100, 50), 4)])
def test_linearize_and_flows_parity(dims, seed):
    c = case(*dims, seed=seed)
    mult = _mult(c, seed)
    m = om.Model(c)
    m.set_directional_mobility(mult)
    prob = orc.with_directional_mobility(orc.problem_from_case(c), mult)
    pv, mean = mixed_state(c, seed=seed)
    st1 = orc.storage(prob, *mixed_state(c, seed=seed + 100))
    m.set_solution(pv, mean)
    m.set_storage1(st1)
    J, r = m.linearize(43200.0, iteration_index=1)
    Jo, ro = orc.linearize(prob, pv, mean, st1, 43200.0, nthreads=16)
    assert np.array_equal(_bits(J), _bits(Jo)) and np.array_equal(_bits(r), _bits(ro))
    fl, fr = m.flows()
    flo, fro = orc.flows(prob, pv, mean)
    assert np.array_equal(_bits(fl), _bits(flo)) and np.array_equal(_bits(fr), _bits(fro))
    # off again: the plain bits
    m.set_directional_mobility(None)
    J0, r0 = m.linearize(43200.0, iteration_index=1)
    Jp, rp = orc.linearize(orc.problem_from_case(c), pv, mean, st1, 43200.0, nthreads=16)
    assert np.array_equal(_bits(J0), _bits(Jp)) and np.array_equal(_bits(r0), _bits(rp))


def test_with_rock_discretization_and_diffusion():
    c = case(12, 9, 5, seed=5)
    mult = _mult(c, 5)
    t, z, vol, thp = random_discretization(c, 5)
    comp = np.random.default_rng(5).uniform(1e-10, 1e-9, c.N)
    m = om.Model(c)
    m.update_discretization(trans=t, depth=z, volume=vol, thpres=thp)
    m.set_rock(comp, None, None)
    m.set_directional_mobility(mult)
    m.set_diffusion(3e-4, 2e-4)
    prob = orc.with_discretization(orc.problem_from_case(c), trans=t, depth=z, volume=vol, thpres=thp)
    prob = orc.with_rock(prob, comp, None, None)
    prob = orc.with_directional_mobility(prob, mult)
    dif = orc.diffusion(3e-4, 2e-4)
    pv, mean = mixed_state(c, seed=5)
    st1 = orc.storage(prob, *mixed_state(c, seed=55))
    m.set_solution(pv, mean)
    m.set_storage1(st1)
    J, r = m.linearize(43200.0, iteration_index=1)
    Jo, ro = orc.linearize(prob, pv, mean, st1, 43200.0, nthreads=16, diffusion=dif)
    assert np.array_equal(_bits(J), _bits(Jo)) and np.array_equal(_bits(r), _bits(ro))


def test_newton_run_parity():
    """Four one-day steps of a 20x16x8 black-oil case with directional
    mobilities at the reference's Newton settings: the oracle's step sizes,
    iteration counts and states bit for bit."""
    c = case(20, 16, 8, seed=9, wells=2)
    mult = np.ones((3, 3, c.N))
    mult[2] = 0.3                  # vertical relative permeabilities a third
    mult[0, 2] = 1.5
    m = om.Model(c)
    m.set_solution(c.pv0, c.mean0)
    m.set_directional_mobility(mult)
    rc, hist, dts, log, st = m.simulate(om.default_params(), 4, 86400.0)
    prob = orc.with_directional_mobility(orc.problem_from_case(c), mult)
    rco, pvo, mo, histo, dtso, logo, _ = orc.simulate(prob, c.pv0, c.mean0, 4, 86400.0, orc.params(), nthreads=8)
    assert rc == 0 and rco == 0
    assert log == logo
    assert np.array_equal(_bits(dts), _bits(dtso))
    assert np.array_equal(_bits(hist), _bits(histo))


def test_invalid_multipliers_are_rejected():
    c = case(3, 3, 2)
    m = om.Model(c)
    bad = np.ones(9 * c.N)
    bad[4] = -1.0
    with pytest.raises(om.OmodelsError):
        m.set_directional_mobility(bad)
