"""CPU tests of the caller-given discretization in the oracle: the
transmissibilities (updateStoredTransmissibilities, tpfalinearizer.hh:868-883),
cell depths, cell volumes and threshold pressures the linearizer takes per
neighbour (tpfalinearizer.hh:443-452).  Pinned by the structured grid's own
values (giving them explicitly changes no bit), by the pattern export, by the
closed forms of the accumulation and source terms, by the threshold-pressure
form of the reference's solvent flux (blackoilsolventmodules.hh:1484-1492) in
closed form and by central differences, and by the volume tie of the upwind
decision, checked against the mobility of the cell it must pick."""
import ctypes as C

import numpy as np
import pytest

import omodels_amd as om
import oracle as orc
from cases import case, mixed_state, own_depth, own_trans


def _nowells(c):
    c.src_cell = c.src_cell[:0].copy()
    c.src_type = c.src_type[:0].copy()
    c.src_a = c.src_a[:0].copy()
    c.src_b = c.src_b[:0].copy()
    c.prob = c.descriptor(om.BlackOilProblem)
    return c


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def test_own_values_given_explicitly_change_no_bit():
    c = case(5, 4, 3, seed=11)
    pv, mean = mixed_state(c, seed=11)
    plain = orc.problem_from_case(c)
    st1 = orc.storage(plain, *mixed_state(c, seed=12))
    J0, r0 = orc.linearize(plain, pv, mean, st1, 86400.0)
    f0, d0 = orc.flows(plain, pv, mean)
    V = c.spec.dx * c.spec.dy * c.spec.dz
    geo = orc.with_discretization(orc.problem_from_case(c), trans=own_trans(plain), depth=own_depth(c),
                                  volume=np.full(c.N, V), thpres=np.zeros(3 * c.N))
    J1, r1 = orc.linearize(geo, pv, mean, st1, 86400.0)
    f1, d1 = orc.flows(geo, pv, mean)
    assert np.array_equal(_bits(J0), _bits(J1)) and np.array_equal(_bits(r0), _bits(r1))
    assert np.array_equal(_bits(f0), _bits(f1)) and np.array_equal(_bits(d0), _bits(d1))
    cells0, rs0, ds0 = orc.source_blocks(plain, pv, mean)
    cells1, rs1, ds1 = orc.source_blocks(geo, pv, mean)
    assert len(cells0) > 0
    assert np.array_equal(_bits(rs0), _bits(rs1)) and np.array_equal(_bits(ds0), _bits(ds1))


def test_given_values_reach_the_pattern():
    c = case(4, 3, 3, seed=5)
    rng = np.random.default_rng(5)
    N = c.N
    t = rng.random(3 * N) * 1e-12
    z = rng.random(N) * 50.0
    prob = orc.with_discretization(orc.problem_from_case(c), trans=t, depth=z)
    pat = orc.pattern(prob)
    for I in range(N):
        for k in range(pat["nb_ptr"][I], pat["nb_ptr"][I + 1]):
            J = pat["nb_cell"][k]
            ax = pat["nb_dir"][k] // 2
            assert pat["nb_trans"][k] == t[ax * N + min(I, J)]
            # tpfalinearizer.hh:450-452: dZg = (zIn - zEx) * gravity
            assert pat["nb_dzg"][k] == (z[I] - z[J]) * prob.gravity


def test_volumes_scale_the_accumulation_and_leave_the_sources():
    """Without face terms (zero transmissibilities) the residual of a cell is
    (storage - storage1) * V_I / dt (tpfalinearizer.hh:776-817) plus its
    source term, whose rate per volume times -V_I is the injected mass."""
    c = case(4, 3, 2, seed=3)
    N = c.N
    rng = np.random.default_rng(3)
    vol = 100.0 + 900.0 * rng.random(N)
    pv, mean = mixed_state(c, seed=3)
    prob = orc.with_discretization(orc.problem_from_case(c), trans=np.zeros(3 * N), volume=vol)
    st0 = orc.storage(prob, pv, mean)
    st1 = orc.storage(prob, *mixed_state(c, seed=4))
    dt = 5000.0
    J, r = orc.linearize(prob, pv, mean, st1, dt)
    src = np.zeros(3 * N)
    cells, rs, _ = orc.source_blocks(prob, pv, mean)
    for k, cell in enumerate(cells):
        src[3 * cell:3 * cell + 3] += rs[3 * k:3 * k + 3]
    want = (st0 - st1) * np.repeat(vol / dt, 3) + src
    assert np.allclose(r, want, rtol=1e-13, atol=0)
    inj = [k for k in range(len(cells)) if c.src_type[k] == 0]
    assert inj
    for k in inj:
        # water mass rate a [kg/s] enters as -a (a / V per volume, times -V)
        assert rs[3 * k + 1] == pytest.approx(-c.src_a[k], rel=1e-15)
    # the diagonal blocks are the storage derivatives times V_I / dt
    plain = orc.with_discretization(orc.problem_from_case(c), trans=np.zeros(3 * N))
    Jp, _ = orc.linearize(plain, pv, mean, st1, dt)
    pat = orc.pattern(prob)
    V0 = c.spec.dx * c.spec.dy * c.spec.dz
    for I in range(N):
        if I in set(cells):
            continue
        o = 9 * pat["diag_off"][I]
        assert np.allclose(J[o:o + 9], Jp[o:o + 9] * (vol[I] / V0), rtol=1e-13, atol=0)


def _flat(nx, ny, nz, seed):
    """a case without wells and without gravity: every phase sees the same
    pressure difference, p_J - p_I"""
    c = _nowells(case(nx, ny, nz, seed=seed))
    pv, mean = mixed_state(c, seed=seed)
    return c, pv, mean


def _dp(pat, pv, k, I):
    return pv[3 * pat["nb_cell"][k] + 1] - pv[3 * I + 1]


def test_threshold_pressure_closed_form():
    """A face with threshold t carries the rates of the pressure difference
    reduced by t towards zero, and none when |dp| <= t."""
    c, pv, mean = _flat(5, 4, 2, seed=21)
    N = c.N
    plain = orc.problem_from_case(c)
    plain.gravity = 0.0
    _, fr0 = orc.flows(plain, pv, mean)
    pat = orc.pattern(plain)
    rng = np.random.default_rng(21)
    # per face: a threshold of 0.3 or 3 times that face's |dp|
    thp = np.zeros(3 * N)
    for I in range(N):
        for k in range(pat["nb_ptr"][I], pat["nb_ptr"][I + 1]):
            J = pat["nb_cell"][k]
            if J > I:
                thp[(pat["nb_dir"][k] // 2) * N + I] = abs(_dp(pat, pv, k, I)) * (0.3 if rng.random() < 0.6 else 3.0)
    prob = orc.with_discretization(orc.problem_from_case(c), thpres=thp)
    prob.gravity = 0.0
    _, fr = orc.flows(prob, pv, mean)
    below = above = 0
    for I in range(N):
        for k in range(pat["nb_ptr"][I], pat["nb_ptr"][I + 1]):
            J = pat["nb_cell"][k]
            t = thp[(pat["nb_dir"][k] // 2) * N + min(I, J)]
            dp = _dp(pat, pv, k, I)
            if abs(dp) > t:
                above += 1
                assert np.allclose(fr[k], fr0[k] * ((abs(dp) - t) / abs(dp)), rtol=1e-12, atol=0)
            else:
                below += 1
                assert np.all(fr[k] == 0.0)
    assert above > 0 and below > 0


def test_threshold_pressure_jacobian_matches_central_differences():
    c, pv, mean = _flat(3, 2, 2, seed=23)
    N = c.N
    base = orc.problem_from_case(c)
    pat = orc.pattern(base)
    thp = np.zeros(3 * N)
    for I in range(N):
        for k in range(pat["nb_ptr"][I], pat["nb_ptr"][I + 1]):
            J = pat["nb_cell"][k]
            if J > I:
                # far from |dp| so that no step crosses it
                f = 0.5 if (I + J) % 3 else 2.0
                thp[(pat["nb_dir"][k] // 2) * N + I] = abs(_dp(pat, pv, k, I)) * f
    prob = orc.with_discretization(orc.problem_from_case(c), thpres=thp)
    prob.gravity = 0.0
    st1 = orc.storage(prob, pv, mean)
    dt = 1e300                             # face terms only
    J, r = orc.linearize(prob, pv, mean, st1, dt)
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


def test_volume_tie_decides_the_upwind_cell():
    """Equal pressures and no gravity: the pressure difference is exactly zero
    and the upwind cell is the larger one (Vin > Vex), the first one for equal
    volumes.  The derivative of the water flux of cell 0 w.r.t. its pressure
    then carries the upwind cell's mobility and 1/Bw:
    A[1][0][water][p] = -invB_w * lam_w * T * rho_w (the up cell's)."""
    c = _nowells(case(2, 1, 1, seed=4))
    pv = np.array([0.2, 200e5, 60.0, 0.6, 200e5, 60.0])
    mean = np.array([1, 1], np.uint8)
    rw = om.default_fluid()[4][4]

    def block10(vol):
        prob = orc.with_discretization(orc.problem_from_case(c), volume=np.array(vol, dtype=float))
        prob.gravity = 0.0
        J, _ = orc.linearize(prob, pv, mean, orc.storage(prob, pv, mean), 1e300)
        pat = orc.pattern(prob)
        T = pat["nb_trans"][0]
        k = pat["nb_off"][0]              # A[1][0], cell 0's face to cell 1
        assert pat["col_idx"][k] == 0 and pat["row_ptr"][1] <= k
        return J[9 * k:9 * k + 9].reshape(3, 3), T, prob

    q = None
    for vol, up in (([1.0, 1.0], 0), ([2.0, 1.0], 0), ([1.0, 2.0], 1)):
        blk, T, prob = block10(vol)
        q = orc.intensive(prob, pv[3 * up:3 * up + 3], mean[up], cell=up)
        want = -q["invB_w"][0] * q["lam_w"][0] * T * rw
        assert blk[1, 1] == pytest.approx(want, rel=1e-13)
    # the two cells' water mobilities differ, so the choice is visible
    q0 = orc.intensive(prob, pv[0:3], 1, cell=0)
    q1 = orc.intensive(prob, pv[3:6], 1, cell=1)
    assert abs(q0["lam_w"][0] - q1["lam_w"][0]) > 0.5 * max(q0["lam_w"][0], q1["lam_w"][0])


@pytest.mark.parametrize("field,value", [("volume", 0.0), ("volume", -1.0), ("face_trans", -1e-13),
                                         ("thpres", -1.0), ("depth", np.nan)])
def test_invalid_discretization_is_rejected(field, value):
    c = case(2, 2, 1)
    N = c.N
    arr = {"volume": np.ones(N), "face_trans": np.zeros(3 * N), "thpres": np.zeros(3 * N),
           "depth": np.zeros(N)}[field]
    arr[1] = value
    kw = {"volume": "volume", "face_trans": "trans", "thpres": "thpres", "depth": "depth"}
    prob = orc.with_discretization(orc.problem_from_case(c), **{kw[field]: arr})
    n = max(prob.n_src, 1)
    cells, r, d = np.zeros(n, np.int32), np.zeros(3 * n), np.zeros(9 * n)
    rc = orc.lib().ob_source_blocks(C.byref(prob), orc.P(c.pv0), orc.P(c.mean0, orc._u8p), orc.P(cells, orc._ip),
                                    orc.P(r), orc.P(d))
    assert rc == -1


def test_two_sources_in_one_cell_add_as_one_rate():
    """problem.source(globI) is one rate per cell: two sources of one cell
    (the corner injector and producer of a 1x1x1 grid) add as the sum of their
    blocks, both of them, not the last one only."""
    c = case(1, 1, 1, seed=1)
    assert len(c.src_cell) == 2 and c.src_cell[0] == c.src_cell[1]
    pv, mean = mixed_state(c, seed=1)
    prob = orc.problem_from_case(c)
    st1 = orc.storage(prob, pv, mean)
    J, r = orc.linearize(prob, pv, mean, st1, 1e300)
    _, rb, db = orc.source_blocks(prob, pv, mean)
    bare = orc.problem_from_case(c)
    bare.n_src = 0
    J0, r0 = orc.linearize(bare, pv, mean, st1, 1e300)
    want = r0 + rb[0:3] + rb[3:6]
    assert np.allclose(r, want, rtol=1e-12, atol=1e-12 * np.abs(rb).max())
    assert np.allclose(J, J0 + db[0:9] + db[9:18], rtol=1e-12, atol=1e-12 * np.abs(db).max())
    for k in range(2):
        assert np.abs(rb[3 * k:3 * k + 3]).max() > 0.0
