"""GPU parity of boundary conditions (omodels_model_set_boundary_conditions;
the boundary terms TpfaLinearizer::linearize_ adds after the cell loop,
tpfalinearizer.hh:849-865, with BlackOilLocalResidualTPFA::
computeBoundaryFlux, blackoillocalresidualtpfa.hh:507-620): residual and
Jacobian bit-exact with the oracle for rate, free and Dirichlet faces, with
the default and with given transmissibilities and face depths, together
with a caller-given discretization; whole Newton runs (eager and replayed
from graphs) bit for bit; removed again, the plain bits."""
import numpy as np
import pytest

import omodels_amd as om
import oracle as orc
from cases import case, mixed_state, random_discretization

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def _boundary_faces(c, seed, frac=0.5):
    """every boundary face of the grid with probability frac, random types,
    rates and states near the cell's own"""
    s = c.spec
    rng = np.random.default_rng(seed)
    pv, mean = mixed_state(c, seed=seed + 50)
    cells, faces = [], []
    for I in range(c.N):
        ijk = (I % s.nx, (I // s.nx) % s.ny, I // (s.nx * s.ny))
        ext = (s.nx, s.ny, s.nz)
        for f in range(6):
            ax = f // 2
            if (ijk[ax] == 0 if f % 2 == 0 else ijk[ax] == ext[ax] - 1) and rng.random() < frac:
                cells.append(I)
                faces.append(f)
    n = len(cells)
    types = rng.integers(0, 4, n)
    rate = rng.uniform(-1e-3, 1e-3, (n, 3))
    state = np.stack([pv[3 * I:3 * I + 3] for I in cells]) if n else np.zeros((0, 3))
    state[:, 1] *= 1.0 + 0.05 * rng.standard_normal(n)
    meaning = mean[cells] if n else np.zeros(0, np.uint8)
    order = rng.permutation(n)             # any input order
    return dict(cell=np.array(cells)[order], face=np.array(faces)[order], type=types[order], rate=rate[order],
                state=state[order], meaning=meaning[order])


@pytest.mark.parametrize("dims,seed", [((1, 1, 1), 1), ((4, 3, 2), 2), ((33, 9, 3), 3), ((64, 64, 16), 4)])
@pytest.mark.parametrize("given", [False, True])
def test_linearize_parity(dims, seed, given):
    c = case(*dims, seed=seed)
    bc = _boundary_faces(c, seed)
    if given:
        rng = np.random.default_rng(seed + 7)
        bc["trans"] = rng.uniform(0.0, 2e-11, len(bc["cell"]))
        bc["zface"] = rng.uniform(0.0, 20.0, len(bc["cell"]))
    print(dims, "boundary faces", len(bc["cell"]), "given", given, flush=True)
    m = om.Model(c)
    m.set_boundary(**bc)
    prob = orc.with_boundary(orc.problem_from_case(c), **bc)
    pv, mean = mixed_state(c, seed=seed)
    st1 = orc.storage(prob, *mixed_state(c, seed=seed + 100))
    m.set_solution(pv, mean)
    m.set_storage1(st1)
    J, r = m.linearize(43200.0, iteration_index=1)
    Jo, ro = orc.linearize(prob, pv, mean, st1, 43200.0, nthreads=16)
    assert np.array_equal(_bits(J), _bits(Jo)) and np.array_equal(_bits(r), _bits(ro))
    # the boundary faces' FLOWS entries
    gc, gf, gfl = m.boundary_flows()
    oc, of, ofl = orc.boundary_flows(prob, pv, mean)
    assert np.array_equal(gc, oc) and np.array_equal(gf, of) and np.array_equal(_bits(gfl), _bits(ofl))
    # removed again: the plain bits
    m.set_boundary(None)
    J0, r0 = m.linearize(43200.0, iteration_index=1)
    Jp, rp = orc.linearize(orc.problem_from_case(c), pv, mean, st1, 43200.0, nthreads=16)
    assert np.array_equal(_bits(J0), _bits(Jp)) and np.array_equal(_bits(r0), _bits(rp))


def test_with_a_given_discretization():
    """depths, volumes and transmissibilities from the caller: the default
    face depths and the cell depth of the gravity term follow them"""
    c = case(12, 9, 5, seed=5)
    t, z, vol, thp = random_discretization(c, 5)
    bc = _boundary_faces(c, 5)
    m = om.Model(c)
    m.update_discretization(trans=t, depth=z, volume=vol, thpres=thp)
    m.set_boundary(**bc)
    prob = orc.with_discretization(orc.problem_from_case(c), trans=t, depth=z, volume=vol, thpres=thp)
    prob = orc.with_boundary(prob, **bc)
    pv, mean = mixed_state(c, seed=5)
    st1 = orc.storage(prob, *mixed_state(c, seed=55))
    m.set_solution(pv, mean)
    m.set_storage1(st1)
    J, r = m.linearize(43200.0, iteration_index=1)
    Jo, ro = orc.linearize(prob, pv, mean, st1, 43200.0, nthreads=16)
    assert np.array_equal(_bits(J), _bits(Jo)) and np.array_equal(_bits(r), _bits(ro))


@pytest.mark.parametrize("graph", [0, 1])
def test_newton_run_parity(graph):
    """Four one-day steps of a 20x16x8 black-oil case whose -x side is held
    at a pressure 5 bar above the initial one (free faces) and whose +x side
    takes water in at a fixed rate: the oracle's step sizes, iteration counts
    and states bit for bit."""
    c = case(20, 16, 8, seed=9, wells=2)
    s = c.spec
    west = [I for I in range(c.N) if I % s.nx == 0]
    east = [I for I in range(c.N) if I % s.nx == s.nx - 1]
    state = np.stack([c.pv0[3 * I:3 * I + 3] for I in west])
    state[:, 1] += 5e5
    rate = np.zeros((len(east), 3))
    rate[:, 1] = -1e-4
    bc = dict(cell=west + east, face=[0] * len(west) + [1] * len(east), type=[2] * len(west) + [1] * len(east),
              rate=np.concatenate([np.zeros((len(west), 3)), rate]), state=np.concatenate([state, np.zeros((len(east), 3))]),
              meaning=np.concatenate([c.mean0[west], np.zeros(len(east), np.uint8)]))
    m = om.Model(c)
    m.set_solution(c.pv0, c.mean0)
    m.set_boundary(**bc)
    rc, hist, dts, log, st = m.simulate(om.default_params(graph=graph), 4, 86400.0)
    prob = orc.with_boundary(orc.problem_from_case(c), **bc)
    rco, pvo, mo, histo, dtso, logo, _ = orc.simulate(prob, c.pv0, c.mean0, 4, 86400.0, orc.params(), nthreads=8)
    print("Newton iterations", st.newton_iters, "linear", st.linear_iters, flush=True)
    assert rc == 0 and rco == 0
    assert log == logo
    assert np.array_equal(_bits(dts), _bits(dtso))
    assert np.array_equal(_bits(hist), _bits(histo))


@pytest.mark.parametrize("what", ["neighbour", "twice", "type", "meaning", "nan_rate", "no_state", "cell"])
def test_invalid_boundaries_are_rejected(what):
    c = case(3, 2, 2)
    kw = dict(cell=[0], face=[0], type=[2], rate=[[0.0, 0.0, 0.0]], state=[c.pv0[:3]], meaning=[c.mean0[0]])
    if what == "neighbour":
        kw["face"] = [1]
    elif what == "twice":
        kw = dict(cell=[0, 0], face=[0, 0], type=[1, 1], rate=np.zeros((2, 3)))
    elif what == "type":
        kw["type"] = [4]
    elif what == "meaning":
        kw["meaning"] = [2]
    elif what == "nan_rate":
        kw["type"] = [1]
        kw["rate"] = [[np.nan, 0.0, 0.0]]
    elif what == "no_state":
        kw["state"] = None
    elif what == "cell":
        kw["cell"] = [c.N]
    m = om.Model(c)
    with pytest.raises(om.OmodelsError):
        m.set_boundary(**kw)
