"""CPU tests of the oracle's boundary conditions (the boundary faces with a
condition that TpfaLinearizer::createMatrix_ collects, tpfalinearizer.hh:
476-497, added after the cell loop at :849-865; the term
BlackOilLocalResidualTPFA::computeBoundaryFlux, blackoillocalresidualtpfa.hh:
507-620, with the boundary gradient of transfluxmodule.hh:324-406).  Pinned
by: type 0 changing no bit; rate faces adding exactly rate x area to the
cell's residual and nothing to the Jacobian; a free face whose state is the
cell's own adding nothing; a free face whose state, transmissibility and
depth are those of a neighbour giving the interior face's flux times the
reference density of each component (the reference's second conversion to
mass); influx when the boundary pressure is above the cell's and outflux
below; shuffled input giving the same bits; central differences of the
Jacobian; the mass balance of a run fed through a rate face; invalid input
rejected."""
import numpy as np
import pytest

import oracle as orc
from cases import case, mixed_state

RHO_COL = lambda prob: np.array([prob.rho_ref[1], prob.rho_ref[0], prob.rho_ref[2]])   # oil, water, gas


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def _lin(prob, pv, mean, st1, dt=43200.0):
    return orc.linearize(prob, pv, mean, st1, dt)


def _diag_block(prob, J, I):
    pat = orc.pattern(prob)
    k = pat["diag_off"][I]
    return J[9 * k:9 * k + 9]


def test_type_none_changes_no_bit():
    c = case(4, 3, 2, seed=3)
    pv, mean = mixed_state(c, seed=3)
    plain = orc.problem_from_case(c)
    st1 = orc.storage(plain, *mixed_state(c, seed=4))
    J0, r0 = _lin(plain, pv, mean, st1)
    prob = orc.with_boundary(orc.problem_from_case(c), cell=[0, 0, 5], face=[0, 2, 4], type=[0, 0, 0])
    J1, r1 = _lin(prob, pv, mean, st1)
    assert np.array_equal(_bits(J0), _bits(J1)) and np.array_equal(_bits(r0), _bits(r1))


def test_rate_faces_add_rate_times_area():
    c = case(4, 3, 2, seed=5)
    pv, mean = mixed_state(c, seed=5)
    plain = orc.problem_from_case(c)
    st1 = orc.storage(plain, *mixed_state(c, seed=6))
    J0, r0 = _lin(plain, pv, mean, st1)
    cells = [0, 3, 3, 23]
    faces = [0, 1, 2, 5]           # -x of cell 0, +x and -y of cell 3, +z of cell 23
    rates = np.array([[1e-3, -2e-3, 5e-5], [-4e-4, 0.0, 1e-6], [2e-4, 3e-4, -7e-7], [0.0, -1e-3, 0.0]])
    prob = orc.with_boundary(orc.problem_from_case(c), cell=cells, face=faces, type=[1] * 4, rate=rates)
    J1, r1 = _lin(prob, pv, mean, st1)
    assert np.array_equal(_bits(J0), _bits(J1))
    s = c.spec
    area = {0: s.dy * s.dz, 1: s.dx * s.dz, 2: s.dx * s.dy}
    want = r0.copy().reshape(-1, 3)
    for I in sorted(set(cells)):
        for b in [b for b in np.argsort(faces, kind="stable") if cells[b] == I]:
            want[I] = want[I] + rates[b] * area[faces[b] // 2]
    assert np.array_equal(_bits(r1), _bits(want.ravel()))


def test_free_face_at_the_cells_own_state_adds_nothing():
    c = case(3, 3, 2, seed=7)
    pv, mean = mixed_state(c, seed=7)
    plain = orc.problem_from_case(c)
    st1 = orc.storage(plain, *mixed_state(c, seed=8))
    J0, r0 = _lin(plain, pv, mean, st1)
    cells = [0, 3, 8]
    faces = [0, 0, 3]              # lateral faces: the face's depth is the cell's
    state = np.stack([pv[3 * I:3 * I + 3] for I in cells])
    for typ in (2, 3):
        prob = orc.with_boundary(orc.problem_from_case(c), cell=cells, face=faces, type=[typ] * 3, state=state,
                                 meaning=mean[cells])
        J1, r1 = _lin(prob, pv, mean, st1)
        assert np.array_equal(_bits(J0), _bits(J1)) and np.array_equal(_bits(r0), _bits(r1))


@pytest.mark.parametrize("seed", [11, 12, 13])
def test_free_face_is_the_interior_flux_times_the_reference_density(seed):
    """Cells 0 and 1 of a 2x1x1 grid share a face; a free condition on the -x
    face of cell 0 with cell 1's state, the shared face's transmissibility
    and cell 0's depth adds the interior face's flux out of cell 0 again,
    each component times its reference density."""
    c = case(2, 1, 1, seed=seed)
    pv, mean = mixed_state(c, seed=seed)
    plain = orc.problem_from_case(c)
    st1 = orc.storage(plain, *mixed_state(c, seed=seed + 1))
    J0, r0 = _lin(plain, pv, mean, st1)
    interior = orc.flows(plain, pv, mean)[0][0]       # the face's flux out of cell 0 (FLOWS)
    pat = orc.pattern(plain)
    T = pat["nb_trans"][pat["nb_ptr"][0]]
    prob = orc.with_boundary(orc.problem_from_case(c), cell=[0], face=[0], type=[2], state=pv[3:6], meaning=mean[1:2],
                             trans=[T])
    J1, r1 = _lin(prob, pv, mean, st1)
    bdy = r1[0:3] - r0[0:3]
    assert np.any(interior != 0.0)
    assert bdy == pytest.approx(RHO_COL(plain) * interior, rel=1e-11, abs=1e-11 * np.max(np.abs(bdy)))
    assert np.array_equal(_bits(r1[3:]), _bits(r0[3:]))


def test_influx_above_and_outflux_below_the_cells_pressure():
    c = case(3, 2, 2, seed=15)
    pv, mean = mixed_state(c, seed=15)
    mean[:] = 1                    # undersaturated: no free gas anywhere
    pv[2::3] = 0.5 * c.pv0[2::3]
    plain = orc.problem_from_case(c)
    st1 = orc.storage(plain, *mixed_state(c, seed=16))
    _, r0 = _lin(plain, pv, mean, st1)
    I = 0
    for dp, sign in ((+5e5, -1.0), (-5e5, +1.0)):
        st = pv[3 * I:3 * I + 3].copy()
        st[1] += dp
        prob = orc.with_boundary(orc.problem_from_case(c), cell=[I], face=[0], type=[2], state=[st], meaning=[1])
        _, r1 = _lin(prob, pv, mean, st1)
        d = (r1 - r0).reshape(-1, 3)
        assert np.all(sign * d[I] >= 0.0) and np.all(sign * d[I, :2] > 0.0), (dp, d[I])
        assert not np.any(np.delete(d, I, axis=0))


def test_shuffled_input_gives_the_same_bits():
    c = case(3, 3, 3, seed=17)
    pv, mean = mixed_state(c, seed=17)
    plain = orc.problem_from_case(c)
    st1 = orc.storage(plain, *mixed_state(c, seed=18))
    rng = np.random.default_rng(17)
    cells = np.array([0, 0, 0, 26, 26, 4, 2, 2])
    faces = np.array([0, 2, 4, 1, 5, 4, 1, 2])
    types = np.array([1, 2, 3, 2, 1, 2, 3, 1])
    rate = rng.uniform(-1e-3, 1e-3, (8, 3))
    state = np.stack([pv[3 * I:3 * I + 3] * [1.0, 1.01, 1.0] for I in cells])
    meaning = mean[cells]
    out = []
    for perm in (np.arange(8), rng.permutation(8)):
        prob = orc.with_boundary(orc.problem_from_case(c), cell=cells[perm], face=faces[perm], type=types[perm],
                                 rate=rate[perm], state=state[perm], meaning=meaning[perm])
        out.append(_lin(prob, pv, mean, st1))
    assert np.array_equal(_bits(out[0][0]), _bits(out[1][0])) and np.array_equal(_bits(out[0][1]), _bits(out[1][1]))


@pytest.mark.parametrize("seed", [1, 2])
def test_jacobian_with_boundary_matches_central_differences(seed):
    c = case(3, 3, 2, seed=seed)
    N = c.N
    pv, mean = mixed_state(c, seed=seed)
    rng = np.random.default_rng(seed)
    cells = [0, 0, 4, 8, 17, 9]
    faces = [0, 4, 4, 1, 5, 0]
    types = [2, 3, 2, 1, 2, 3]
    state = np.stack([pv[3 * I:3 * I + 3] * [1.0, 1.0 + 0.05 * rng.standard_normal(), 1.0] for I in cells])
    rate = rng.uniform(-1e-3, 1e-3, (6, 3))
    prob = orc.with_boundary(orc.problem_from_case(c), cell=cells, face=faces, type=types, rate=rate, state=state,
                             meaning=mean[cells])
    st1 = orc.storage(prob, *mixed_state(c, seed=seed + 10))
    dt = 86400.0
    J, r = orc.linearize(prob, pv, mean, st1, dt)
    pat = orc.pattern(prob)
    steps = [1e-7, 1e-2, 1e-7]
    for col in range(N):
        for var in range(3):
            h = steps[var] if not (var == 2 and mean[col] == 1) else 1e-5
            pp = pv.copy(); pp[3 * col + var] += h
            pm = pv.copy(); pm[3 * col + var] -= h
            _, rp = orc.linearize(prob, pp, mean, st1, dt)
            _, rm = orc.linearize(prob, pm, mean, st1, dt)
            fd = (rp - rm) / (2 * h)
            for row in range(N):
                ks = np.where(pat["col_idx"][pat["row_ptr"][row]:pat["row_ptr"][row + 1]] == col)[0]
                blk = J[9 * (pat["row_ptr"][row] + ks[0]):][:9].reshape(3, 3) if len(ks) else np.zeros((3, 3))
                scale = np.max(np.abs(J[9 * pat["row_ptr"][row]:9 * pat["row_ptr"][row + 1]]).reshape(-1, 3, 3)[:, :, var])
                assert np.max(np.abs(blk[:, var] - fd[3 * row:3 * row + 3])) <= 2e-5 * max(scale, 1e-300) + 1e-9


def test_mass_balance_of_a_run_fed_through_rate_faces():
    """Water injected through the -x faces of the first column for four
    days of a closed 6x4x3 grid: the water in place grows by the injected
    mass, oil and gas in place stay, to the Newton tolerance."""
    c = case(6, 4, 3, seed=21)
    s = c.spec
    col = [I for I in range(c.N) if I % s.nx == 0]
    q = 2e-4                                       # kg/(s m^2), into the cell: negative out-rate
    rate = np.zeros((len(col), 3))
    rate[:, 1] = -q
    prob = orc.problem_from_case(c)
    prob.n_src = 0                                 # no wells: closed but for the rate faces
    prob = orc.with_boundary(prob, cell=col, face=[0] * len(col), type=[1] * len(col), rate=rate)
    V = s.dx * s.dy * s.dz
    m0 = orc.storage(prob, c.pv0, c.mean0).reshape(-1, 3).sum(axis=0) * V
    rc, pv, mean, hist, dts, log, _ = orc.simulate(prob, c.pv0, c.mean0, 4, 86400.0, orc.params(newton_tol=1e-9),
                                                   nthreads=8)
    assert rc == 0
    m1 = orc.storage(prob, pv, mean).reshape(-1, 3).sum(axis=0) * V
    injected = q * len(col) * s.dy * s.dz * float(np.sum(dts))
    assert m1[1] - m0[1] == pytest.approx(injected, rel=1e-6)
    assert abs(m1[0] - m0[0]) <= 1e-6 * injected and abs(m1[2] - m0[2]) <= 1e-6 * injected


@pytest.mark.parametrize("what", ["neighbour", "twice", "type", "meaning", "nan_rate", "no_state", "face"])
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
    elif what == "face":
        kw["face"] = [6]
    prob = orc.with_boundary(orc.problem_from_case(c), **kw)
    with pytest.raises(AssertionError):
        orc.flows(prob, c.pv0, c.mean0)


def test_boundary_flows_are_the_terms_added_to_the_residual():
    """One face per cell: each face's FLOWS entry is exactly what its cell's
    residual gains; type-0 faces give zeros; the faces come in cell and face
    order whatever the input order."""
    c = case(4, 4, 2, seed=23)
    pv, mean = mixed_state(c, seed=23)
    plain = orc.problem_from_case(c)
    st1 = orc.storage(plain, *mixed_state(c, seed=24))
    _, r0 = _lin(plain, pv, mean, st1)
    cells = np.array([15, 0, 5, 10, 3])
    faces = np.array([3, 0, 4, 4, 1])
    types = np.array([2, 1, 3, 0, 2])
    rng = np.random.default_rng(23)
    state = np.stack([pv[3 * I:3 * I + 3] * [1.0, 1.03, 1.0] for I in cells])
    prob = orc.with_boundary(orc.problem_from_case(c), cell=cells, face=faces, type=types,
                             rate=rng.uniform(-1e-3, 1e-3, (5, 3)), state=state, meaning=mean[cells])
    _, r1 = _lin(prob, pv, mean, st1)
    bc, bf, fl = orc.boundary_flows(prob, pv, mean)
    assert list(bc) == sorted(cells) and len(bf) == 5
    for k in range(5):
        I = bc[k]
        assert np.array_equal(_bits(r1[3 * I:3 * I + 3]), _bits(r0[3 * I:3 * I + 3] + fl[k]))
        if types[list(cells).index(I)] == 0:
            assert not np.any(fl[k])
