"""The assembly's z march over several planes per workgroup (kChunk >= 2:
the below / current / above columns carried in registers from plane to
plane) against the oracle, bit for bit.  Small grids get one plane per
workgroup; the march runs at the headline's 100x100x50 (kChunk 2, 25
workgroups along z), at 200x200x50 (kChunk 5) and in the ensemble, whose
smaller launch target gives a 64x64x40 member kChunk 2 and a 100x100x50
member kChunk 5.  Every kernel variant (plain, diffusion, dispersion, both,
caller-given discretization) runs at kChunk >= 2."""
import numpy as np
import pytest

import omodels_amd as om
import oracle as orc
from cases import case, mixed_state, random_discretization

pytestmark = pytest.mark.gpu

D_OIL, D_GAS = 3e-4, 2e-4


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def _k_chunk(nx, ny, nz, target):
    gx, gy = (nx + 31) // 32, (ny + 7) // 8
    gz = min(max((target + gx * gy - 1) // (gx * gy), 1), nz)
    return (nz + gz - 1) // gz


@pytest.mark.parametrize("dims,terms", [((100, 100, 50), t) for t in ("plain", "diffusion", "dispersion", "both",
                                                                        "given")]
                         + [((200, 200, 50), "plain"), ((200, 200, 50), "given")])
def test_linearize_parity_with_several_planes_per_workgroup(dims, terms):
    assert _k_chunk(*dims, 2048) >= 2
    c = case(*dims, seed=17)
    pv, mean = mixed_state(c, seed=17)
    prob = orc.problem_from_case(c)
    st1 = orc.storage(prob, *mixed_state(c, seed=117))
    m = om.Model(c)
    m.set_solution(pv, mean)
    m.set_storage1(st1)
    dif = dsp = None
    if terms in ("diffusion", "both"):
        m.set_diffusion(D_OIL, D_GAS)
        dif = orc.diffusion(D_OIL, D_GAS)
    if terms in ("dispersion", "both"):
        E = np.random.default_rng(17).uniform(5.0, 5e3, c.N)
        m.set_dispersion(cell_dispersion=E)
        dsp = orc.dispersion(cell_dispersion=E)
    if terms == "given":
        t, z, vol, thp = random_discretization(c, 17)
        m.update_discretization(trans=t, depth=z, volume=vol, thpres=thp)
        prob = orc.with_discretization(prob, trans=t, depth=z, volume=vol, thpres=thp)
    J, r = m.linearize(43200.0, iteration_index=1)
    if dsp is None:
        Jo, ro = orc.linearize(prob, pv, mean, st1, 43200.0, nthreads=16, diffusion=dif)
    else:
        Jo, ro, nvo = orc.linearize(prob, pv, mean, st1, 43200.0, nthreads=16, diffusion=dif, dispersion=dsp)
        assert np.array_equal(_bits(m.cell_velocity()), _bits(nvo))
    assert np.array_equal(_bits(J), _bits(Jo))
    assert np.array_equal(_bits(r), _bits(ro))


def test_ensemble_members_with_several_planes_per_workgroup():
    """one step of two members whose assembly marches 2 and 5 planes per
    workgroup, against the oracle run of each alone, bit for bit"""
    ms = [case(64, 64, 40, seed=41, wells=2), case(100, 100, 50, seed=42, wells=1)]
    assert _k_chunk(64, 64, 40, 512) >= 2 and _k_chunk(100, 100, 50, 512) >= 5
    prm = om.default_params()
    srv = om.Server(prm, 64e9, 16)
    ids = [srv.submit(c, 1, 86400.0, 0.0) for c in ms]
    srv.run()
    for c, i in zip(ms, ids):
        r = srv.result(i)
        assert r["status"] == 0
        rco, pvo, mo, histo, dtso, logo, tim = orc.simulate(orc.problem_from_case(c), c.pv0, c.mean0, 1, 86400.0,
                                                            orc.params(), nthreads=16)
        assert rco == 0
        assert np.array_equal(r["dts"], dtso)
        assert np.array_equal(r["mean"], mo)
        assert np.array_equal(_bits(r["pv"]), _bits(pvo))
