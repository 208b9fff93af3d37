"""GPU parity of the caller-given discretization
(omodels_linearizer_update_discretization): transmissibilities
(updateStoredTransmissibilities, tpfalinearizer.hh:868-883), cell depths,
cell volumes and threshold pressures (the per-neighbour inputs of
tpfalinearizer.hh:443-452) on distorted grids.  Bar: residual, Jacobian,
FLOWS, source blocks and the pattern's exported values bit-exact with the
oracle; simulated states bit-exact over the steps run."""
import numpy as np
import pytest

import omodels_amd as om
import oracle as orc
from cases import case, mixed_state, own_depth, own_trans, random_discretization

pytestmark = pytest.mark.gpu

D_OIL, D_GAS = 3e-4, 2e-4


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def _geo_of(prob):
    return prob._geo_keep_args


def _both(c, seed):
    t, z, vol, thp = random_discretization(c, seed)
    m = om.Model(c)
    m.update_discretization(trans=t, depth=z, volume=vol, thpres=thp)
    prob = orc.with_discretization(orc.problem_from_case(c), trans=t, depth=z, volume=vol, thpres=thp)
    prob._geo_keep_args = (t, z, vol, thp)
    return m, prob


@pytest.mark.parametrize("dims,seed", [((1, 1, 1), 1), ((3, 2, 1), 2), ((4, 3, 2), 3), ((33, 9, 3), 4),
                                       ((40, 40, 10), 5)])
def test_linearize_flows_and_sources_parity(dims, seed):
    c = case(*dims, seed=seed)
    pv, mean = mixed_state(c, seed=seed)
    m, prob = _both(c, seed)
    st1 = orc.storage(prob, *mixed_state(c, seed=seed + 100))
    m.set_solution(pv, mean)
    m.set_storage1(st1)
    J, r = m.linearize(43200.0, iteration_index=1)
    Jo, ro = orc.linearize(prob, pv, mean, st1, 43200.0)
    assert np.array_equal(_bits(J), _bits(Jo))
    assert np.array_equal(_bits(r), _bits(ro))
    fl, fr = m.flows()
    flo, fro = orc.flows(prob, pv, mean)
    assert np.array_equal(_bits(fl), _bits(flo)) and np.array_equal(_bits(fr), _bits(fro))
    cg, rg, dg = m.source_blocks()
    co, ro_, do = orc.source_blocks(prob, pv, mean)
    assert np.array_equal(cg, co)
    assert np.array_equal(_bits(rg), _bits(ro_)) and np.array_equal(_bits(dg), _bits(do))
    # the separate sparse source path: the system without sources, then each
    # source's block added in the list's order
    m.set_separate_sparse_sources(True)
    J2, r2 = m.linearize(43200.0, iteration_index=1)
    bare = orc.with_discretization(orc.problem_from_case(c), *_geo_of(prob))
    bare.n_src = 0
    Jb, rb = orc.linearize(bare, pv, mean, st1, 43200.0)
    diag = orc.pattern(prob)["diag_off"]
    for k, cell in enumerate(co):
        rb[3 * cell:3 * cell + 3] += ro_[3 * k:3 * k + 3]
        Jb[9 * diag[cell]:9 * diag[cell] + 9] += do[9 * k:9 * k + 9]
    assert np.array_equal(_bits(J2), _bits(Jb)) and np.array_equal(_bits(r2), _bits(rb))
    if len(set(co.tolist())) == len(co):
        assert np.array_equal(_bits(J2), _bits(Jo)) and np.array_equal(_bits(r2), _bits(ro))
    # the pattern reports the given values
    pg, po = m.pattern(), orc.pattern(prob)
    assert np.array_equal(_bits(pg["nb_trans"]), _bits(po["nb_trans"]))
    assert np.array_equal(_bits(pg["nb_dzg"]), _bits(po["nb_dzg"]))


@pytest.mark.parametrize("terms", ["diffusion", "dispersion", "both"])
@pytest.mark.parametrize("dims,seed", [((4, 3, 2), 2), ((33, 9, 3), 3)])
def test_with_diffusion_and_dispersion_parity(dims, seed, terms):
    """the given discretization in the kernels with the diffusive and the
    dispersive face terms, over two linearizations (velocities handed on)"""
    c = case(*dims, seed=seed)
    pv, mean = mixed_state(c, seed=seed)
    m, prob = _both(c, seed + 10)
    st1 = orc.storage(prob, *mixed_state(c, seed=seed + 100))
    m.set_solution(pv, mean)
    m.set_storage1(st1)
    E = np.random.default_rng(seed).uniform(5.0, 5e3, c.N)
    dif = dsp = None
    if terms in ("diffusion", "both"):
        m.set_diffusion(D_OIL, D_GAS)
        dif = orc.diffusion(D_OIL, D_GAS)
    if terms in ("dispersion", "both"):
        m.set_dispersion(cell_dispersion=E)
        dsp = orc.dispersion(cell_dispersion=E)
    nv = np.zeros(c.N)
    for it in range(2):
        J, r = m.linearize(43200.0, iteration_index=1)
        if dsp is None:
            Jo, ro = orc.linearize(prob, pv, mean, st1, 43200.0, diffusion=dif)
        else:
            Jo, ro, nv_next = orc.linearize(prob, pv, mean, st1, 43200.0, diffusion=dif, dispersion=dsp, nv=nv)
        assert np.array_equal(_bits(J), _bits(Jo)), it
        assert np.array_equal(_bits(r), _bits(ro)), it
        if dsp is not None:
            nv = nv_next
            assert np.array_equal(_bits(m.cell_velocity()), _bits(nv)), it
    fl, fr = m.flows()
    flo, fro = orc.flows(prob, pv, mean, diffusion=dif, dispersion=dsp, nv=nv if dsp is not None else None)
    assert np.array_equal(_bits(fl), _bits(flo)) and np.array_equal(_bits(fr), _bits(fro))


def test_own_values_through_the_given_path_change_no_bit():
    """The grid's own values given explicitly run the kernels' given-geometry
    form and reproduce the plain form's bits; all NULL returns to the plain
    form after a distorted grid."""
    c = case(33, 9, 3, seed=6)
    pv, mean = mixed_state(c, seed=6)
    prob = orc.problem_from_case(c)
    st1 = orc.storage(prob, *mixed_state(c, seed=106))
    V = c.spec.dx * c.spec.dy * c.spec.dz
    plain = om.Model(c)
    plain.set_solution(pv, mean)
    plain.set_storage1(st1)
    J0, r0 = plain.linearize(43200.0, iteration_index=1)
    fl0, fr0 = plain.flows()
    m = om.Model(c)
    m.set_solution(pv, mean)
    m.set_storage1(st1)
    m.update_discretization(trans=own_trans(prob), depth=own_depth(c), volume=np.full(c.N, V),
                            thpres=np.zeros(3 * c.N))
    J1, r1 = m.linearize(43200.0, iteration_index=1)
    fl1, fr1 = m.flows()
    assert np.array_equal(_bits(J0), _bits(J1)) and np.array_equal(_bits(r0), _bits(r1))
    assert np.array_equal(_bits(fl0), _bits(fl1)) and np.array_equal(_bits(fr0), _bits(fr1))
    t, z, vol, thp = random_discretization(c, 6)
    m.update_discretization(trans=t, depth=z, volume=vol, thpres=thp)
    J2, _ = m.linearize(43200.0, iteration_index=1)
    assert not np.array_equal(_bits(J2), _bits(J0))
    m.update_discretization()
    J3, r3 = m.linearize(43200.0, iteration_index=1)
    assert np.array_equal(_bits(J0), _bits(J3)) and np.array_equal(_bits(r0), _bits(r3))
    assert np.array_equal(_bits(m.pattern()["nb_dzg"]), _bits(plain.pattern()["nb_dzg"]))


def test_simulate_parity_on_a_distorted_grid():
    """Newton with time-step control on a distorted grid with threshold
    pressures: the oracle's step sizes, iteration counts and states, bit for
    bit, at the reference's settings."""
    c = case(20, 20, 5, seed=31)
    m, prob = _both(c, 31)
    m.set_solution(c.pv0, c.mean0)
    rc, hist, dts, log, st = m.simulate(om.default_params(), 4, 86400.0)
    rco, pvo, mo, histo, dtso, logo, tim = orc.simulate(prob, c.pv0, c.mean0, 4, 86400.0, orc.params(), nthreads=4)
    assert rc == 0 and rco == 0
    assert np.array_equal(dts, dtso)
    assert log == logo
    assert np.array_equal(_bits(hist), _bits(histo))


@pytest.mark.parametrize("field,value", [("volume", 0.0), ("volume", np.inf), ("trans", -1.0), ("thpres", -1.0),
                                         ("depth", np.nan)])
def test_invalid_discretization_is_rejected(field, value):
    c = case(3, 2, 2)
    n = 3 * c.N if field in ("trans", "thpres") else c.N
    a = np.ones(n)
    a[2] = value
    m = om.Model(c)
    with pytest.raises(om.OmodelsError):
        m.update_discretization(**{field: a})
    with pytest.raises(ValueError):
        m.update_discretization(**{field: np.ones(n + 1)})
