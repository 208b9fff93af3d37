"""GPU parity of the sub-domain linearization
(omodels_linearizer_linearize_subdomain; linearizeDomain(domain),
tpfalinearizer.hh:226-245) with the oracle, bit for bit: two sub-domains in a
row over a previous full linearization (rows outside keep their values but
for the column blocks added into them), with sources, boundary conditions,
rock terms and directional mobilities, and with a caller-given
discretization; all cells give the full linearization; the storage cache is
not refreshed; with separate sparse source terms every well's terms are
added; invalid domains are rejected."""
import numpy as np
import pytest

import omodels_amd as om
import oracle as orc
from cases import case, mixed_state, random_discretization

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def _extras(c, seed):
    rng = np.random.default_rng(seed)
    comp = rng.uniform(1e-10, 1e-9, c.N)
    mult = rng.uniform(0.2, 3.0, (3, 3, c.N))
    s = c.spec
    west = [I for I in range(c.N) if I % s.nx == 0]
    rate = np.zeros((len(west), 3))
    rate[:, 1] = -1e-4
    bc = dict(cell=west, face=[0] * len(west), type=[1] * len(west), rate=rate)
    return comp, mult, bc


@pytest.mark.parametrize("dims,seed,given", [((6, 5, 4), 1, False), ((33, 9, 3), 2, False), ((40, 40, 10), 3, False),
                                             ((12, 9, 5), 4, True)])
def test_two_subdomains_after_a_full_linearization(dims, seed, given):
    c = case(*dims, seed=seed, wells=1)
    comp, mult, bc = _extras(c, seed)
    m = om.Model(c)
    prob = orc.problem_from_case(c)
    if given:
        t, z, vol, thp = random_discretization(c, seed)
        m.update_discretization(trans=t, depth=z, volume=vol, thpres=thp)
        prob = orc.with_discretization(prob, trans=t, depth=z, volume=vol, thpres=thp)
    m.set_rock(comp, None, None)
    m.set_directional_mobility(mult)
    m.set_boundary(**bc)
    prob = orc.with_boundary(orc.with_directional_mobility(orc.with_rock(prob, comp, None, None), mult), **bc)
    st1 = orc.storage(prob, *mixed_state(c, seed=seed + 9))
    pv, mean = mixed_state(c, seed=seed)
    m.set_solution(pv, mean)
    m.set_storage1(st1)
    J, r = m.linearize(43200.0, iteration_index=1)
    Jo, ro = orc.linearize(prob, pv, mean, st1, 43200.0, nthreads=16)
    assert np.array_equal(_bits(J), _bits(Jo))
    rng = np.random.default_rng(seed)
    for rep in range(2):
        pv, mean = mixed_state(c, seed=seed + 20 + rep)
        cells = rng.choice(c.N, max(1, c.N // 3), replace=False)
        m.set_solution(pv, mean)
        J, r = m.linearize_subdomain(43200.0, cells, iteration_index=0)
        Jo, ro = orc.linearize_domain(prob, pv, mean, st1, 43200.0, cells, Jo, ro)
        print(dims, "sub-domain", rep, "cells", len(cells), flush=True)
        assert np.array_equal(_bits(J), _bits(Jo)) and np.array_equal(_bits(r), _bits(ro))
    assert np.array_equal(_bits(m.get_storage1()), _bits(st1))     # not refreshed on a sub-domain
    # all cells: the full linearization
    J, r = m.linearize_subdomain(43200.0, rng.permutation(c.N))
    Jf, rf = orc.linearize(prob, pv, mean, st1, 43200.0, nthreads=16)
    assert np.array_equal(_bits(J), _bits(Jf)) and np.array_equal(_bits(r), _bits(rf))


def test_separate_sparse_sources_add_every_well():
    c = case(10, 8, 3, seed=5, wells=2)
    prob = orc.problem_from_case(c)
    m = om.Model(c)
    m.set_separate_sparse_sources(True)
    st1 = orc.storage(prob, *mixed_state(c, seed=6))
    pv, mean = mixed_state(c, seed=5)
    m.set_solution(pv, mean)
    m.set_storage1(st1)
    J, r = m.linearize(43200.0)
    Jo, ro = orc.linearize(prob, pv, mean, st1, 43200.0)
    cells = np.arange(0, c.N, 3)
    pv2, mean2 = mixed_state(c, seed=7)
    m.set_solution(pv2, mean2)
    J, r = m.linearize_subdomain(43200.0, cells)
    # the oracle: the domain without sources, then every source's block added in order
    nosrc = orc.problem_from_case(c)
    nosrc.n_src = 0
    Jo, ro = orc.linearize_domain(nosrc, pv2, mean2, st1, 43200.0, cells, Jo, ro)
    sc, sr, sd = orc.source_blocks(prob, pv2, mean2)
    pat = orc.pattern(prob)
    for k in range(len(sc)):
        I = sc[k]
        ro[3 * I:3 * I + 3] = ro[3 * I:3 * I + 3] + sr[3 * k:3 * k + 3]
        d = pat["diag_off"][I]
        Jo[9 * d:9 * d + 9] = Jo[9 * d:9 * d + 9] + sd[9 * k:9 * k + 9]
    assert np.array_equal(_bits(J), _bits(Jo)) and np.array_equal(_bits(r), _bits(ro))


@pytest.mark.parametrize("what", ["empty", "twice", "range", "diffusion"])
def test_invalid_subdomains_are_rejected(what):
    c = case(4, 3, 2)
    m = om.Model(c)
    m.set_solution(c.pv0, c.mean0)
    cells = {"empty": [], "twice": [1, 1], "range": [c.N], "diffusion": [0, 1]}[what]
    if what == "diffusion":
        m.set_diffusion(1e-9, 1e-9)
    with pytest.raises(om.OmodelsError):
        m.linearize_subdomain(86400.0, cells)
