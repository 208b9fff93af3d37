"""CPU tests that pin the oracle (the CPU restatement of the reference path)
by mathematics the reference fixes, independently of the GPU library:
brute-force topology, closed forms of the tables, finite-difference
Jacobians, conservation and hydrostatic invariants, the defining property of
ILU(0), exact solves on systems where the answer is known, and the Newton
chopping/switching and time-step rules stated by the reference."""
import numpy as np
import pytest

import omodels_amd as om
import oracle as orc
from cases import case, mixed_state, rel_err


def _case_nowells(nx, ny, nz, seed=3, **kw):
    c = case(nx, ny, nz, seed=seed, **kw)
    c.src_cell = c.src_cell[:0].copy()
    c.src_type = c.src_type[:0].copy()
    c.src_a = c.src_a[:0].copy()
    c.src_b = c.src_b[:0].copy()
    c.prob = c.descriptor(om.BlackOilProblem)
    return c


# ----------------------------------------------------------------------------
# A1: topology by brute force
@pytest.mark.parametrize("dims", [(1, 1, 1), (3, 1, 1), (1, 4, 1), (2, 2, 2), (4, 3, 2), (5, 1, 3)])
def test_pattern_brute_force(dims):
    nx, ny, nz = dims
    c = case(nx, ny, nz)
    pat = orc.pattern(orc.problem_from_case(c))
    N = nx * ny * nz
    idx = lambda i, j, k: i + nx * (j + ny * k)
    rows = []
    nbs = []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                nb = []
                # intersection order of a structured grid: -x,+x,-y,+y,-z,+z
                for (di, dj, dk) in [(-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)]:
                    a, b, cc = i + di, j + dj, k + dk
                    if 0 <= a < nx and 0 <= b < ny and 0 <= cc < nz:
                        nb.append(idx(a, b, cc))
                nbs.append(nb)
                rows.append(sorted(nb + [idx(i, j, k)]))
    rp = np.cumsum([0] + [len(r) for r in rows])
    assert np.array_equal(pat["row_ptr"], rp)
    assert np.array_equal(pat["col_idx"], np.concatenate(rows))
    for I in range(N):
        assert pat["col_idx"][pat["diag_off"][I]] == I
        got = pat["nb_cell"][pat["nb_ptr"][I]:pat["nb_ptr"][I + 1]]
        assert list(got) == nbs[I]
        for k, J in zip(range(pat["nb_ptr"][I], pat["nb_ptr"][I + 1]), got):
            off = pat["nb_off"][k]
            assert pat["col_idx"][off] == I and rp[J] <= off < rp[J + 1]


def test_transmissibility_series_law_and_depth_terms():
    c = case(3, 2, 2, seed=8)
    prob = orc.problem_from_case(c)
    pat = orc.pattern(prob)
    s = c.spec
    area = {0: s.dy * s.dz, 1: s.dx * s.dz, 2: s.dx * s.dy}
    h = {0: s.dx, 1: s.dy, 2: s.dz}
    N = c.N
    for I in range(N):
        for k in range(pat["nb_ptr"][I], pat["nb_ptr"][I + 1]):
            J = pat["nb_cell"][k]
            ax = pat["nb_dir"][k] // 2
            K1, K2 = c.perm[ax * N + I], c.perm[ax * N + J]
            # two half cells in series: A / (h/2/K1 + h/2/K2)
            T = area[ax] / (0.5 * h[ax] / K1 + 0.5 * h[ax] / K2)
            assert abs(pat["nb_trans"][k] - T) <= 1e-13 * T
            dz = (J // (s.nx * s.ny) - I // (s.nx * s.ny)) * s.dz
            assert abs(pat["nb_dzg"][k] - (-dz) * -9.81) <= 1e-12
            assert pat["nb_area"][k] == area[ax]


# ----------------------------------------------------------------------------
# A4: PVT tables and intensive quantities
def _fluid():
    oil, no, gas, ng, consts = om.default_fluid()
    return oil.reshape(4, no), gas.reshape(3, ng), consts


def test_tables_reproduce_nodes_and_midpoints():
    oil, gas, consts = _fluid()
    c = case(1, 1, 1)
    prob = orc.problem_from_case(c)
    for i in range(oil.shape[1]):
        p = oil[0, i]
        sat = orc.intensive(prob, [0.2, p, 0.1], 0)          # meaning Sg: saturated oil
        assert sat["invB_o"][0] == pytest.approx(oil[2, i], rel=1e-15, abs=0)
        assert sat["Rs"][0] == pytest.approx(oil[1, i], rel=1e-15, abs=0)
        # undersaturated oil at its own bubble point equals the saturated oil
        und = orc.intensive(prob, [0.2, p, oil[1, i]], 1)
        assert und["invB_o"][0] == pytest.approx(oil[2, i], rel=1e-12)
    for i in range(oil.shape[1] - 1):
        pm = 0.5 * (oil[0, i] + oil[0, i + 1])
        sat = orc.intensive(prob, [0.2, pm, 0.1], 0)
        assert sat["invB_o"][0] == pytest.approx(0.5 * (oil[2, i] + oil[2, i + 1]), rel=1e-13)
        assert sat["Rs"][0] == pytest.approx(0.5 * (oil[1, i] + oil[1, i + 1]), rel=1e-13)
        # slope of the table is the pressure derivative
        slope = (oil[2, i + 1] - oil[2, i]) / (oil[0, i + 1] - oil[0, i])
        assert sat["invB_o"][2] == pytest.approx(slope, rel=1e-13)
    for i in range(gas.shape[1]):
        g = orc.intensive(prob, [0.2, gas[0, i], 0.3], 0)
        assert g["invB_g"][0] == pytest.approx(gas[1, i], rel=1e-15)


def test_densities_mobilities_and_water_compressibility():
    oil, gas, consts = _fluid()
    c = case(1, 1, 1)
    prob = orc.problem_from_case(c)
    co, pref, cw, muw, rw, ro, rg = consts
    for pv, mean in [([0.25, 180e5, 0.15], 0), ([0.3, 250e5, 120.0], 1), ([0.0, 90e5, 0.0], 0)]:
        q = orc.intensive(prob, pv, mean)
        Rs = q["Rs"][0]
        # oil phase density: mass of stock-tank oil plus dissolved gas per reservoir volume
        assert q["rho_o"][0] == pytest.approx(q["invB_o"][0] * (ro + Rs * rg), rel=1e-14)
        assert q["rho_w"][0] == pytest.approx(q["invB_w"][0] * rw, rel=1e-15)
        # linear relative permeability: lambda_w = Sw / mu_w
        assert q["lam_w"][0] == pytest.approx(pv[0] / muw, rel=1e-15)
        X = cw * (pv[1] - pref)
        assert abs(q["invB_w"][0] - np.exp(X)) <= abs(X) ** 3
        So = 1.0 - pv[0] - (pv[2] if mean == 0 else 0.0)
        assert q["So"][0] == pytest.approx(So, abs=1e-15)


@pytest.mark.parametrize("pv,mean", [([0.27, 173e5, 0.11], 0), ([0.31, 227e5, 101.3], 1), ([0.05, 64e5, 0.4], 0)])
def test_intensive_derivatives_match_finite_differences(pv, mean):
    prob = orc.problem_from_case(case(1, 1, 1))
    base = orc.intensive(prob, pv, mean)
    steps = [1e-6, 1.0, 1e-6 if mean == 0 else 1e-4]
    for var in range(3):
        hp = list(pv)
        hm = list(pv)
        hp[var] += steps[var]
        hm[var] -= steps[var]
        qp = orc.intensive(prob, hp, mean)
        qm = orc.intensive(prob, hm, mean)
        for k in base:
            fd = (qp[k][0] - qm[k][0]) / (2 * steps[var])
            ad = base[k][1 + var]
            scale = max(abs(ad), abs(base[k][0]) / abs(pv[var] if pv[var] else 1.0) * 1e-6, 1e-30)
            assert abs(fd - ad) <= 1e-5 * scale + 1e-12 * abs(base[k][0]), (k, var, fd, ad)


# ----------------------------------------------------------------------------
# A2-A7: residual and Jacobian
def _lin(c, pv, mean, st1, dt):
    return orc.linearize(orc.problem_from_case(c), pv, mean, st1, dt)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_jacobian_matches_central_differences(seed):
    c = case(3, 3, 2, seed=seed)
    pv, mean = mixed_state(c, seed=seed)
    prob = orc.problem_from_case(c)
    st1 = orc.storage(prob, *mixed_state(c, seed=seed + 10))
    dt = 86400.0
    J, r = _lin(c, pv, mean, st1, dt)
    pat = orc.pattern(prob)
    N = c.N
    steps = [1e-7, 1e-2, 1e-7]
    for col in range(N):
        for var in range(3):
            h = steps[var] if not (var == 2 and mean[col] == 1) else 1e-5
            pp = pv.copy(); pp[3 * col + var] += h
            pm = pv.copy(); pm[3 * col + var] -= h
            _, rp = _lin(c, pp, mean, st1, dt)
            _, rm = _lin(c, pm, mean, st1, dt)
            fd = (rp - rm) / (2 * h)
            # compare with column `col`, variable `var` of the BCRS Jacobian
            for row in range(N):
                ks = np.where(pat["col_idx"][pat["row_ptr"][row]:pat["row_ptr"][row + 1]] == col)[0]
                blk = J[9 * (pat["row_ptr"][row] + ks[0]):][:9].reshape(3, 3) if len(ks) else np.zeros((3, 3))
                ad = blk[:, var]
                fdr = fd[3 * row:3 * row + 3]
                scale = np.max(np.abs(J[9 * pat["row_ptr"][row]:9 * pat["row_ptr"][row + 1]]).reshape(-1, 3, 3)[:, :, var])
                assert np.max(np.abs(ad - fdr)) <= 2e-5 * max(scale, 1e-300) + 1e-9, (row, col, var, ad, fdr)


def test_interior_fluxes_cancel_without_sources():
    c = _case_nowells(5, 4, 3, seed=5)
    pv, mean = mixed_state(c, seed=5)
    prob = orc.problem_from_case(c)
    st1 = orc.storage(prob, pv, mean)          # zero accumulation
    J, r = _lin(c, pv, mean, st1, 1e30)        # and a vanishing accumulation derivative
    R = r.reshape(-1, 3)
    assert np.max(np.abs(R.sum(axis=0))) <= 1e-12 * np.abs(R).sum(axis=0).max()
    # every face contributes equal and opposite derivatives to A[J][I] and A[I][I]:
    # the column sums of dF/dx_I over all rows vanish without accumulation
    pat = orc.pattern(prob)
    colsum = np.zeros((c.N, 3, 3))
    for I in range(c.N):
        for k in range(pat["row_ptr"][I], pat["row_ptr"][I + 1]):
            colsum[pat["col_idx"][k]] += J[9 * k:9 * k + 9].reshape(3, 3)
    scale = np.abs(J).max()
    assert np.abs(colsum).max() <= 1e-12 * scale


def test_hydrostatic_oil_column_has_no_flux():
    c = _case_nowells(1, 1, 12, seed=2, sw_init=0.0)
    prob = orc.problem_from_case(c)
    st1 = orc.storage(prob, c.pv0, c.mean0)
    J, r = _lin(c, c.pv0, c.mean0, st1, 86400.0)
    # gravity flux scale of one face if the column were not in equilibrium
    pat = orc.pattern(prob)
    T = pat["nb_trans"].max()
    q = orc.intensive(prob, c.pv0[:3], c.mean0[0])
    scale = T * q["lam_o"][0] * q["rho_o"][0] * 9.81 * c.spec.dz * 786.0
    assert np.abs(r).max() <= 1e-8 * scale


def test_accumulation_term_closed_form():
    c = _case_nowells(1, 1, 1)
    pv, mean = np.array([0.3, 210e5, 0.2]), np.array([0], np.uint8)
    prob = orc.problem_from_case(c)
    st0 = orc.storage(prob, pv, mean)
    old = np.array([0.25, 200e5, 0.1])
    st1 = orc.storage(prob, old, mean)
    dt = 1000.0
    J, r = _lin(c, pv, mean, st1, dt)
    V = c.spec.dx * c.spec.dy * c.spec.dz
    assert np.allclose(r, (st0 - st1) * V / dt, rtol=1e-14, atol=0)
    q = orc.intensive(prob, pv, 0)
    phi = c.poro[0]
    rw, ro, rg = om.default_fluid()[4][4:7]
    # water mass per bulk volume: phi Sw invBw rho_w
    assert st0[1] == pytest.approx(phi * pv[0] * q["invB_w"][0] * rw, rel=1e-14)
    So = 1 - pv[0] - pv[2]
    assert st0[2] == pytest.approx(phi * (pv[2] * q["invB_g"][0] + So * q["invB_o"][0] * q["Rs"][0]) * rg, rel=1e-13)


# ----------------------------------------------------------------------------
# Sparse source terms: the blocks added after the cell loop
def _source_setup(seed):
    c = case(6, 5, 3, seed=seed)
    pv, mean = mixed_state(c, seed=seed)
    prob = orc.problem_from_case(c)
    st1 = orc.storage(prob, *mixed_state(c, seed=seed + 20))
    return c, pv, mean, prob, st1


@pytest.mark.parametrize("seed", [4, 5])
def test_source_blocks_are_what_the_cell_loop_adds(seed):
    c, pv, mean, prob, st1 = _source_setup(seed)
    J, r = orc.linearize(prob, pv, mean, st1, 86400.0)
    bare = orc.problem_from_case(c)
    bare.n_src = 0
    J0, r0 = orc.linearize(bare, pv, mean, st1, 86400.0)
    cells, rs, ds = orc.source_blocks(prob, pv, mean)
    assert np.array_equal(cells, c.src_cell)
    # the producers are open, so their blocks carry derivatives
    assert all(np.abs(ds[9 * s:9 * s + 9]).max() > 0 for s in range(len(cells)) if c.src_type[s] == 1)
    diag = orc.pattern(prob)["diag_off"]
    for s, I in enumerate(cells):
        r0[3 * I:3 * I + 3] += rs[3 * s:3 * s + 3]
        J0[9 * diag[I]:9 * diag[I] + 9] += ds[9 * s:9 * s + 9]
    assert np.array_equal(r, r0)
    assert np.array_equal(J, J0)


def test_source_blocks_closed_form_and_derivatives():
    c, pv, mean, prob, st1 = _source_setup(4)
    cells, rs, ds = orc.source_blocks(prob, pv, mean)
    rw, ro, rg = om.default_fluid()[4][4:7]
    for s, I in enumerate(cells):
        R, D = rs[3 * s:3 * s + 3], ds[9 * s:9 * s + 9].reshape(3, 3)
        a, b = c.src_a[s], c.src_b[s]
        if c.src_type[s] == 0:
            # water injection at a fixed rate: -rate in the water equation, no derivatives
            assert R[1] == pytest.approx(-a, rel=1e-15) and R[0] == 0.0 and R[2] == 0.0
            assert not D.any()
            continue
        # producer at fixed bottom-hole pressure: (rate of each phase at surface) * density,
        # the gas equation also takes the oil's dissolved gas
        q = orc.intensive(prob, pv[3 * I:3 * I + 3], mean[I], cell=I)
        drive = q["p"][0] - b
        assert drive > 0
        w = a * drive * q["lam_w"][0] * q["invB_w"][0] * rw
        o = a * drive * q["lam_o"][0] * q["invB_o"][0] * ro
        g = a * drive * (q["lam_g"][0] * q["invB_g"][0] + q["Rs"][0] * q["lam_o"][0] * q["invB_o"][0]) * rg
        assert R == pytest.approx([o, w, g], rel=1e-13)
        # derivatives against central differences of the blocks themselves
        for var, h in enumerate([1e-7, 1e-2, 1e-7 if mean[I] == 0 else 1e-5]):
            pp, pm = pv.copy(), pv.copy()
            pp[3 * I + var] += h
            pm[3 * I + var] -= h
            fd = (orc.source_blocks(prob, pp, mean)[1] - orc.source_blocks(prob, pm, mean)[1])[3 * s:3 * s + 3] / (2 * h)
            assert np.abs(D[:, var] - fd).max() <= 2e-6 * np.abs(D).max(), (s, var, D[:, var], fd)


# ----------------------------------------------------------------------------
# Face rates for output (FLOWS / FLORES)
def test_flows_sum_to_the_residual_in_order():
    """Without sources or accumulation the residual of a cell is the sum of its
    face fluxes in intersection order: the FLOWS are exactly those terms."""
    c = _case_nowells(5, 4, 3, seed=5)
    pv, mean = mixed_state(c, seed=5)
    prob = orc.problem_from_case(c)
    J, r = orc.linearize(prob, pv, mean, orc.storage(prob, pv, mean), 1e300)
    fl, _ = orc.flows(prob, pv, mean)
    nbp = orc.pattern(prob)["nb_ptr"]
    acc = np.zeros(3 * c.N)
    for I in range(c.N):
        a = np.zeros(3)
        for k in range(nbp[I], nbp[I + 1]):
            a = a + fl[k]
        acc[3 * I:3 * I + 3] = a
    assert np.array_equal(acc.view(np.uint64), r.view(np.uint64))


def test_flores_closed_form_and_antisymmetry():
    c = case(4, 3, 3, seed=9)
    pv, mean = mixed_state(c, seed=9)
    prob = orc.problem_from_case(c)
    fl, fr = orc.flows(prob, pv, mean)
    pat = orc.pattern(prob)
    nbp, nbc = pat["nb_ptr"], pat["nb_cell"]
    q = [orc.intensive(prob, pv[3 * I:3 * I + 3], mean[I], cell=I) for I in range(c.N)]
    comp = {"w": 1, "o": 0, "g": 2}
    checked = 0
    for I in range(c.N):
        for k in range(nbp[I], nbp[I + 1]):
            Jc = nbc[k]
            T, A, dzg = pat["nb_trans"][k], pat["nb_area"][k], pat["nb_dzg"][k]
            for ph in "wog":
                lamI, lamJ = q[I]["lam_" + ph][0], q[Jc]["lam_" + ph][0]
                want = 0.0
                if lamI > 0 or lamJ > 0:
                    rho = 0.5 * (q[I]["rho_" + ph][0] + q[Jc]["rho_" + ph][0])
                    dp = rho * dzg + q[Jc]["p"][0] - q[I]["p"][0]
                    up = lamJ if dp > 0 else lamI
                    want = dp * up * (-T / A) * A
                assert fr[k, comp[ph]] == pytest.approx(want, rel=1e-12, abs=1e-300)
                checked += want != 0.0
            # the same face seen from the neighbour carries the opposite rates
            kb = [kk for kk in range(nbp[Jc], nbp[Jc + 1]) if nbc[kk] == I][0]
            for arr in (fl, fr):
                scale = np.abs(arr[k]).max() + 1e-300
                assert np.abs(arr[k] + arr[kb]).max() <= 1e-10 * scale
    assert checked > 0


# ----------------------------------------------------------------------------
# A11-A12: linear algebra
def _system(dims, seed=1):
    c = case(*dims, seed=seed)
    pv, mean = mixed_state(c, seed=seed)
    prob = orc.problem_from_case(c)
    st1 = orc.storage(prob, *mixed_state(c, seed=seed + 50))
    J, r = orc.linearize(prob, pv, mean, st1, 86400.0)
    return c, J, r, orc.pattern(prob)


def _dense(c, J, pat):
    n = 3 * c.N
    A = np.zeros((n, n))
    rp, ci = pat["row_ptr"], pat["col_idx"]
    for i in range(c.N):
        for k in range(rp[i], rp[i + 1]):
            A[3 * i:3 * i + 3, 3 * ci[k]:3 * ci[k] + 3] = J[9 * k:9 * k + 9].reshape(3, 3)
    return A


@pytest.mark.parametrize("dims", [(4, 3, 2), (6, 1, 3), (3, 3, 3)])
def test_ilu0_factors_satisfy_lu_equals_a_on_the_pattern(dims):
    c, J, r, pat = _system(dims)
    rc, lu = orc.ilu0(c.N, 3, pat["row_ptr"], pat["col_idx"], J)
    assert rc == 0
    n = c.N
    rp, ci, dg = pat["row_ptr"], pat["col_idx"], pat["diag_off"]
    L = np.zeros((3 * n, 3 * n))
    U = np.zeros((3 * n, 3 * n))
    for i in range(n):
        L[3 * i:3 * i + 3, 3 * i:3 * i + 3] = np.eye(3)
        for k in range(rp[i], rp[i + 1]):
            j = ci[k]
            B = lu[9 * k:9 * k + 9].reshape(3, 3)
            if j < i:
                L[3 * i:3 * i + 3, 3 * j:3 * j + 3] = B
            elif j == i:
                U[3 * i:3 * i + 3, 3 * i:3 * i + 3] = np.linalg.inv(B)   # stored inverted
            else:
                U[3 * i:3 * i + 3, 3 * j:3 * j + 3] = B
    A = _dense(c, J, pat)
    LU = L @ U
    mask = np.zeros_like(A, dtype=bool)
    for i in range(n):
        for k in range(rp[i], rp[i + 1]):
            mask[3 * i:3 * i + 3, 3 * ci[k]:3 * ci[k] + 3] = True
    scale = np.abs(A).max()
    assert np.abs((LU - A)[mask]).max() <= 1e-12 * scale
    # applying the preconditioner solves with L U exactly
    d = np.random.default_rng(1).standard_normal(3 * n)
    v = orc.ilu0_apply(n, 3, rp, ci, lu, d)
    assert np.allclose(LU @ v, d, rtol=0, atol=1e-10 * np.abs(d).max())


def test_ilu0_is_exact_lu_for_block_tridiagonal():
    c, J, r, pat = _system((9, 1, 1), seed=2)
    rc, lu = orc.ilu0(c.N, 3, pat["row_ptr"], pat["col_idx"], J)
    A = _dense(c, J, pat)
    v = orc.ilu0_apply(c.N, 3, pat["row_ptr"], pat["col_idx"], lu, r)
    x = np.linalg.solve(A, r)
    assert np.allclose(v, x, rtol=1e-9, atol=1e-12 * np.abs(x).max())


@pytest.mark.parametrize("precond", [0, 1])
def test_bicgstab_matches_dense_solve(precond):
    """ILU(0), and block Jacobi with three sweeps (one or two sweeps do not
    precondition this matrix well enough to converge)."""
    c, J, r, pat = _system((5, 4, 2), seed=3)
    order = 3 if precond == 1 else 0
    rc, x, its = orc.bicgstab(c.N, 3, pat["row_ptr"], pat["col_idx"], J, r,
                              orc.params(lin_tol=1e-13, precond=precond, prec_order=order, lin_max_iter=5000,
                                         lin_abs_tol=0.0))
    assert rc == 0
    xd = np.linalg.solve(_dense(c, J, pat), r)
    assert np.max(np.abs(x - xd)) <= 1e-8 * np.abs(xd).max()
    # at the reference tolerance the max-norm residual is reduced by 1e-3
    rc, x, its = orc.bicgstab(c.N, 3, pat["row_ptr"], pat["col_idx"], J, r,
                              orc.params(precond=precond, prec_order=order))
    res = r - orc.spmv(c.N, 3, pat["row_ptr"], pat["col_idx"], J, x)
    assert rc == 0 and np.abs(res).max() <= 1.2e-3 * np.abs(r).max()


def test_float_solver_scalar():
    c, J, r, pat = _system((5, 4, 2), seed=3)
    rc, x, its = orc.bicgstab(c.N, 3, pat["row_ptr"], pat["col_idx"], J, r,
                              orc.params(lin_tol=1e-5, solver_float=1))
    xd = np.linalg.solve(_dense(c, J, pat), r)
    assert rc == 0
    assert np.max(np.abs(x - xd)) <= 1e-3 * np.abs(xd).max()


# ----------------------------------------------------------------------------
# A16: Newton update, chopping and switching
def _update(c, cur, mean, delta, was=None, **kw):
    prob = orc.problem_from_case(c)
    was = np.zeros(c.N, np.uint8) if was is None else was
    return orc.update_primary(prob, cur, mean, delta, was, orc.params(**kw))


def test_pressure_and_saturation_chopping():
    c = _case_nowells(2, 1, 1)
    cur = np.array([0.3, 200e5, 0.2, 0.3, 200e5, 0.2])
    mean = np.array([0, 0], np.uint8)
    delta = np.array([0.5, -1e8, 0.1, 0.05, 1e5, 0.05])
    rc, nxt, nm, ws, nsw = _update(c, cur, mean, delta)
    # |dp| <= dpMaxRel * p
    assert nxt[1] == pytest.approx(200e5 * 1.3)
    assert nxt[4] == pytest.approx(200e5 - 1e5)
    # saturation deltas scaled so that the largest |dS| (here dSo = -0.6) is dsMax
    alpha = 0.2 / 0.6
    assert nxt[0] == pytest.approx(0.3 - 0.5 * alpha)
    assert nxt[2] == pytest.approx(0.2 - 0.1 * alpha)
    assert nxt[3] == pytest.approx(0.25) and nxt[5] == pytest.approx(0.15)


def test_gas_appears_and_disappears_at_the_bubble_point():
    c = _case_nowells(1, 1, 1)
    oil, gas, consts = _fluid()
    rs_sat = lambda p: np.interp(p, oil[0], oil[1])
    # undersaturated oil whose pressure falls below its bubble point: free gas appears
    p0 = 200e5
    cur = np.array([0.2, p0, rs_sat(p0)])
    rc, nxt, nm, ws, nsw = _update(c, cur, np.array([1], np.uint8), np.array([0.0, 5e5, 0.0]))
    assert nm[0] == 0 and nxt[2] == 0.0 and nsw == 1
    # free gas that disappears (Sg < 0) switches back to dissolved gas at Rs_sat(p)
    cur = np.array([0.2, p0, 0.01])
    rc, nxt, nm, ws, nsw = _update(c, cur, np.array([0], np.uint8), np.array([0.0, 0.0, 0.02]))
    assert nm[0] == 1 and nxt[2] == pytest.approx(rs_sat(p0), rel=1e-12) and nsw == 1
    # hysteresis: a cell switched in the last iteration needs Sg < -1e-5 to switch again
    cur = np.array([0.2, p0, 0.0])
    rc, nxt, nm, ws, nsw = _update(c, cur, np.array([0], np.uint8), np.array([0.0, 0.0, 5e-6]),
                                   was=np.array([1], np.uint8))
    assert nm[0] == 0 and nsw == 0
    rc, nxt, nm, ws, nsw = _update(c, cur, np.array([0], np.uint8), np.array([0.0, 0.0, 5e-6]))
    assert nm[0] == 1 and nsw == 1


def test_non_finite_update_is_a_numerical_problem():
    c = _case_nowells(2, 1, 1)
    cur = np.array([0.3, 200e5, 0.2, 0.3, 200e5, 0.2])
    rc, *_ = _update(c, cur, np.zeros(2, np.uint8), np.array([0, np.nan, 0, 0, 0, 0.0]))
    assert rc == 2


# ----------------------------------------------------------------------------
# A15/A18: Newton loop, time-step control, conservation, failure recovery
def test_converged_state_and_water_mass_balance():
    """A converged step satisfies its own residual to the Newton tolerance, and
    with an injector only, the water mass gained equals the injected mass."""
    c = case(6, 5, 3, seed=4)
    keep = c.src_type == 0
    c.src_cell, c.src_type = c.src_cell[keep].copy(), c.src_type[keep].copy()
    c.src_a, c.src_b = c.src_a[keep].copy(), c.src_b[keep].copy()
    c.prob = c.descriptor(om.BlackOilProblem)
    prob = orc.problem_from_case(c)
    tol = 1e-9
    dt = 86400.0
    rc, pv, mean, hist, dts, log, tim = orc.simulate(prob, c.pv0, c.mean0, 1, dt, orc.params(newton_tol=tol))
    assert rc == 0 and dts[0] == dt
    st0 = orc.storage(prob, c.pv0, c.mean0)
    J, r = orc.linearize(prob, pv, mean, st0, dt)
    assert np.abs(r).max() <= tol
    V = c.spec.dx * c.spec.dy * c.spec.dz
    dM = (orc.storage(prob, pv, mean) - st0).reshape(-1, 3).sum(axis=0) * V
    injected = c.src_a.sum() * dt
    assert abs(dM[1] - injected) <= c.N * tol * dt
    assert abs(dM[0]) <= c.N * tol * dt and abs(dM[2]) <= c.N * tol * dt


def test_time_step_control_follows_iteration_counts():
    c = case(6, 5, 3, seed=4)
    prob = orc.problem_from_case(c)
    rc, pv, mean, hist, dts, log, tim = orc.simulate(prob, c.pv0, c.mean0, 5, 86400.0, orc.params())
    assert rc == 0
    for k in range(4):
        n = sum(1 for l in log if l[0] == k and l[1] == 0)
        expect = dts[k] * (1 + (10 - n) / 10 / 1.2) if n <= 10 else dts[k] / (1 + (n - 10) / 10)
        assert dts[k + 1] == pytest.approx(expect, rel=1e-14)


def test_failed_newton_halves_the_step_and_restores_the_state():
    c = case(6, 5, 3, seed=4)
    prob = orc.problem_from_case(c)
    # one linear iteration cannot reach the linear tolerance: every attempt fails
    rc, pv, mean, hist, dts, log, tim = orc.simulate(prob, c.pv0, c.mean0, 1, 86400.0,
                                                     orc.params(lin_max_iter=1, max_dt_divisions=3))
    assert rc == 3
    attempts = sorted({(l[1], l[7]) for l in log})
    assert [a for a, _ in attempts] == [0, 1, 2]
    assert [d for _, d in attempts] == [86400.0, 43200.0, 21600.0]
    # the state handed back is the start of the step
    assert np.array_equal(pv, c.pv0) and np.array_equal(mean, c.mean0)


def test_oracle_assembly_is_independent_of_thread_count():
    c = case(7, 6, 5, seed=6)
    pv, mean = mixed_state(c, seed=6)
    prob = orc.problem_from_case(c)
    st1 = orc.storage(prob, pv, mean)
    J1, r1 = orc.linearize(prob, pv, mean, st1, 1e4, nthreads=1)
    J4, r4 = orc.linearize(prob, pv, mean, st1, 1e4, nthreads=4)
    assert np.array_equal(J1, J4) and np.array_equal(r1, r4)


# ----------------------------------------------------------------------------
# Molecular diffusion in the oil phase (blackoildiffusionmodule.hh:168-238,
# face term of blackoillocalresidualtpfa.hh:469-479)
D_OIL, D_GAS = 3e-4, 2e-4           # large, so that the term is far above round-off


def _diff(**kw):
    return orc.diffusion(kw.pop("oil_in_oil", D_OIL), kw.pop("gas_in_oil", D_GAS), **kw)


def _no_acc(prob, pv, mean, dif):
    """residual and Jacobian of the face terms only (no accumulation, no wells)"""
    return orc.linearize(prob, pv, mean, orc.storage(prob, pv, mean), 1e300, diffusion=dif)


def _two_cells(rs_i, rs_j, sw=0.2, p=200e5):
    c = _case_nowells(2, 1, 1, seed=4)
    prob = orc.problem_from_case(c)
    prob.gravity = 0.0                # equal pressures, no gravity: no advective flux
    pv = np.array([sw, p, rs_i, sw, p, rs_j])
    return c, prob, pv, np.array([1, 1], np.uint8)


def test_diffusion_closed_form_between_two_cells():
    """Two undersaturated cells at one pressure: the only flux is diffusion.
    Gas leaves the cell with the larger Rs, oil enters it, in the ratio of the
    coefficients; the amounts follow from the mean So/Bo, the Rs difference
    over (rho_o/rho_g + mean Rs) and the harmonic mean of A phi / (h/2)."""
    c, prob, pv, mean = _two_cells(140.0, 90.0)
    _, r0 = _no_acc(prob, pv, mean, None)
    assert np.all(r0 == 0.0)
    _, r = _no_acc(prob, pv, mean, _diff())
    qi = orc.intensive(prob, pv[:3], 1, cell=0)
    qj = orc.intensive(prob, pv[3:], 1, cell=1)
    so = 1.0 - 0.2
    bs = 0.5 * (so * qi["invB_o"][0] + so * qj["invB_o"][0])
    rho_w, rho_o, rho_g = om.default_fluid()[4][4:7]
    conv = 1.0 / (rho_o / rho_g + 0.5 * (140.0 + 90.0))
    A, h = c.spec.dy * c.spec.dz, c.spec.dx
    d0, d1 = A * c.poro[0] / (h / 2), A * c.poro[1] / (h / 2)
    dif = d0 * d1 / (d0 + d1)
    gas = bs * (140.0 - 90.0) * conv * dif * D_GAS
    assert gas > 0.0
    assert r[2] == pytest.approx(gas, rel=1e-13)                      # gas component (index 2)
    assert r[0] == pytest.approx(-gas * D_OIL / D_GAS, rel=1e-13)     # oil component (index 0)
    assert r[1] == 0.0                                                # no water term
    assert np.array_equal(r[3:], -r[:3])                              # the same face from J


def test_diffusion_mole_fraction_form():
    """With molar masses the conversion factor uses rho_o M_g / (rho_g M_o)."""
    c, prob, pv, mean = _two_cells(140.0, 90.0)
    _, rm = _no_acc(prob, pv, mean, _diff())
    _, rx = _no_acc(prob, pv, mean, _diff(mw_oil=0.2, mw_gas=0.016))
    rho_w, rho_o, rho_g = om.default_fluid()[4][4:7]
    rs = 0.5 * (140.0 + 90.0)
    ratio = (rho_o / rho_g + rs) / (rho_o * 0.016 / (rho_g * 0.2) + rs)
    assert rx[2] == pytest.approx(rm[2] * ratio, rel=1e-13)
    assert rx[0] == pytest.approx(rm[0] * ratio, rel=1e-13)


def test_diffusion_needs_oil():
    """Without oil in either cell (mean So/Bo below 1e-6) the term vanishes."""
    c, prob, pv, mean = _two_cells(0.0, 0.0)
    pv = np.array([0.3, 200e5, 0.7, 0.4, 200e5, 0.6])
    mean = np.array([0, 0], np.uint8)
    J0, r0 = _no_acc(prob, pv, mean, None)
    J1, r1 = _no_acc(prob, pv, mean, _diff())
    assert np.array_equal(r0.view(np.uint64), r1.view(np.uint64))
    assert np.array_equal(J0.view(np.uint64), J1.view(np.uint64))


def test_diffusion_face_diffusivity_given_or_derived():
    """The derived face diffusivities equal an explicit array of the same
    values bit for bit, and the term is linear in them."""
    c = _case_nowells(4, 3, 2, seed=6)
    pv, mean = mixed_state(c, seed=6)
    prob = orc.problem_from_case(c)
    N = c.N
    nx, ny, nz = 4, 3, 2
    h = (c.spec.dx, c.spec.dy, c.spec.dz)
    area = (c.spec.dy * c.spec.dz, c.spec.dx * c.spec.dz, c.spec.dx * c.spec.dy)
    fd = np.zeros(3 * N)
    for I in range(N):
        ijk = (I % nx, (I // nx) % ny, I // (nx * ny))
        for ax, (n, stride) in enumerate(((nx, 1), (ny, nx), (nz, nx * ny))):
            if ijk[ax] + 1 < n:
                d0 = area[ax] * c.poro[I] * (0.5 * h[ax]) / (0.5 * h[ax]) ** 2
                d1 = area[ax] * c.poro[I + stride] * (0.5 * h[ax]) / (0.5 * h[ax]) ** 2
                fd[ax * N + I] = d0 * d1 / (d0 + d1)
    _, r0 = _no_acc(prob, pv, mean, None)
    Jd, rd = _no_acc(prob, pv, mean, _diff())
    Jg, rg = _no_acc(prob, pv, mean, _diff(face_diffusivity=fd))
    assert np.array_equal(rd.view(np.uint64), rg.view(np.uint64))
    assert np.array_equal(Jd.view(np.uint64), Jg.view(np.uint64))
    _, r2 = _no_acc(prob, pv, mean, _diff(face_diffusivity=2.0 * fd))
    part = rd - r0
    assert np.abs(part).max() > 1e-6 * np.abs(r0).max()
    assert rel_err(r2 - r0, 2.0 * part) <= 1e-9


def test_diffusion_jacobian_matches_central_differences():
    """Derivatives of the diffusive part (with and without minus without)
    against central differences of the residual, all cells and variables."""
    c = _case_nowells(3, 2, 2, seed=8)
    pv, mean = mixed_state(c, seed=8)
    prob = orc.problem_from_case(c)
    st1 = orc.storage(prob, pv, mean)
    dif = _diff()

    def part(x):
        Jd, rd = orc.linearize(prob, x, mean, st1, 1e300, diffusion=dif)
        J0, r0 = orc.linearize(prob, x, mean, st1, 1e300)
        return Jd - J0, rd - r0

    Jp, rp_ = part(pv)
    pat = orc.pattern(prob)
    N = c.N
    steps = [1e-6, 1.0, 1e-6]
    scale = np.abs(Jp).reshape(-1, 3, 3).max(axis=(0, 1))
    assert scale.min() > 0.0
    for col in range(N):
        for var in range(3):
            h = steps[var] if not (var == 2 and mean[col] == 1) else 1e-4
            pp = pv.copy(); pp[3 * col + var] += h
            pm = pv.copy(); pm[3 * col + var] -= h
            fd = (part(pp)[1] - part(pm)[1]) / (2 * h)
            for row in range(N):
                ks = np.where(pat["col_idx"][pat["row_ptr"][row]:pat["row_ptr"][row + 1]] == col)[0]
                blk = Jp[9 * (pat["row_ptr"][row] + ks[0]):][:9].reshape(3, 3) if len(ks) else np.zeros((3, 3))
                assert np.max(np.abs(blk[:, var] - fd[3 * row:3 * row + 3])) <= 1e-5 * scale[var], (row, col, var)


def test_diffusion_conserves_and_flows_sum_to_the_residual():
    """With diffusion the interior face terms still cancel over the domain,
    and the FLOWS (which include it) add up to the residual bit for bit."""
    c = _case_nowells(5, 4, 3, seed=5)
    pv, mean = mixed_state(c, seed=5)
    prob = orc.problem_from_case(c)
    dif = _diff()
    J, r = _no_acc(prob, pv, mean, dif)
    _, r0 = _no_acc(prob, pv, mean, None)
    R, P = r.reshape(-1, 3), (r - r0).reshape(-1, 3)
    assert np.abs(P).max() > 0.0
    assert np.max(np.abs(R.sum(axis=0))) <= 1e-12 * np.abs(R).sum(axis=0).max()
    fl, _ = orc.flows(prob, pv, mean, diffusion=dif)
    nbp = orc.pattern(prob)["nb_ptr"]
    acc = np.zeros(3 * c.N)
    for I in range(c.N):
        a = np.zeros(3)
        for k in range(nbp[I], nbp[I + 1]):
            a = a + fl[k]
        acc[3 * I:3 * I + 3] = a
    assert np.array_equal(acc.view(np.uint64), r.view(np.uint64))


# ----------------------------------------------------------------------------
# Dispersion in the oil phase (blackoildispersionmodule.hh:176-254), driven by
# each cell's oil velocity of the previous linearization
def _disp_lin(prob, pv, mean, nv, dsp, dif=None):
    return orc.linearize(prob, pv, mean, orc.storage(prob, pv, mean), 1e300, diffusion=dif, dispersion=dsp, nv=nv)


def test_dispersion_closed_form_between_two_cells():
    """Two undersaturated cells at one pressure: with velocities v0, v1 of the
    previous linearization the gas flux is mean(1/Bo) mean(v) (Rs_I - Rs_J)
    / (rho_o/rho_g + mean Rs) times the face dispersivity, the oil flux its
    negative; zero velocities (the first linearization) give no flux."""
    c, prob, pv, mean = _two_cells(140.0, 90.0)
    E = np.array([12.0, 30.0])
    dsp = orc.dispersion(cell_dispersion=E)
    J0, r0, nv0 = _disp_lin(prob, pv, mean, np.zeros(2), dsp)
    assert np.all(r0 == 0.0) and np.all(nv0 == 0.0)        # no flow, no velocity
    _, r, _ = _disp_lin(prob, pv, mean, np.array([2e-6, 5e-6]), dsp)
    qi = orc.intensive(prob, pv[:3], 1, cell=0)
    qj = orc.intensive(prob, pv[3:], 1, cell=1)
    b = 0.5 * (qi["invB_o"][0] + qj["invB_o"][0])
    rho_w, rho_o, rho_g = om.default_fluid()[4][4:7]
    conv = 1.0 / (rho_o / rho_g + 0.5 * (140.0 + 90.0))
    A, h = c.spec.dy * c.spec.dz, c.spec.dx
    e0, e1 = A * E[0] / (h / 2), A * E[1] / (h / 2)
    gas = b * 0.5 * (2e-6 + 5e-6) * conv * (e0 * e1 / (e0 + e1)) * (140.0 - 90.0)
    assert gas > 0.0
    assert r[2] == pytest.approx(gas, rel=1e-13)
    assert r[0] == pytest.approx(-gas, rel=1e-13)
    assert r[1] == 0.0
    assert np.array_equal(r[3:], -r[:3])


def test_dispersion_velocity_is_the_largest_oil_darcy_velocity():
    """The velocity a linearization leaves for the next one is, per cell, the
    largest |FLORES oil rate / face area| over its faces."""
    c = _case_nowells(5, 4, 3, seed=11)
    pv, mean = mixed_state(c, seed=11)
    prob = orc.problem_from_case(c)
    dsp = orc.dispersion(cell_dispersion=np.full(c.N, 10.0))
    _, _, nv = _disp_lin(prob, pv, mean, np.zeros(c.N), dsp)
    _, fr = orc.flows(prob, pv, mean)
    pat = orc.pattern(prob)
    for I in range(c.N):
        ks = range(pat["nb_ptr"][I], pat["nb_ptr"][I + 1])
        want = max([abs(fr[k, 0] / pat["nb_area"][k]) for k in ks], default=0.0)
        assert nv[I] == want
    assert nv.max() > 0.0


def test_dispersion_given_or_derived_and_jacobian():
    c = _case_nowells(3, 2, 2, seed=13)
    pv, mean = mixed_state(c, seed=13)
    prob = orc.problem_from_case(c)
    N, nx, ny, nz = c.N, 3, 2, 2
    E = np.random.default_rng(13).uniform(5.0, 50.0, N)
    h = (c.spec.dx, c.spec.dy, c.spec.dz)
    area = (c.spec.dy * c.spec.dz, c.spec.dx * c.spec.dz, c.spec.dx * c.spec.dy)
    fd = np.zeros(3 * N)
    for I in range(N):
        ijk = (I % nx, (I // nx) % ny, I // (nx * ny))
        for ax, (n, stride) in enumerate(((nx, 1), (ny, nx), (nz, nx * ny))):
            if ijk[ax] + 1 < n:
                e0 = area[ax] * E[I] * (0.5 * h[ax]) / (0.5 * h[ax]) ** 2
                e1 = area[ax] * E[I + stride] * (0.5 * h[ax]) / (0.5 * h[ax]) ** 2
                fd[ax * N + I] = e0 * e1 / (e0 + e1)
    nv = np.random.default_rng(14).uniform(1e-7, 1e-5, N)
    Jd, rd, _ = _disp_lin(prob, pv, mean, nv, orc.dispersion(cell_dispersion=E))
    Jg, rg, _ = _disp_lin(prob, pv, mean, nv, orc.dispersion(face_dispersivity=fd))
    assert np.array_equal(rd.view(np.uint64), rg.view(np.uint64))
    assert np.array_equal(Jd.view(np.uint64), Jg.view(np.uint64))
    # derivatives of the dispersive part against central differences
    st1 = orc.storage(prob, pv, mean)
    dsp = orc.dispersion(face_dispersivity=1e4 * fd)

    def part(x):
        J1, r1, _ = orc.linearize(prob, x, mean, st1, 1e300, dispersion=dsp, nv=nv)
        J0, r0 = orc.linearize(prob, x, mean, st1, 1e300)
        return J1 - J0, r1 - r0

    Jp, _ = part(pv)
    pat = orc.pattern(prob)
    scale = np.abs(Jp).reshape(-1, 3, 3).max(axis=(0, 1))
    assert scale[1] > 0.0 and scale[2] > 0.0
    # round-off floor of the differences of the full residuals
    rmax = np.abs(orc.linearize(prob, pv, mean, st1, 1e300)[1]).max()
    steps = [1e-6, 1.0, 1e-6]
    for col in range(N):
        for var in range(3):
            hh = steps[var] if not (var == 2 and mean[col] == 1) else 1e-4
            pp = pv.copy(); pp[3 * col + var] += hh
            pm = pv.copy(); pm[3 * col + var] -= hh
            fdr = (part(pp)[1] - part(pm)[1]) / (2 * hh)
            for row in range(N):
                ks = np.where(pat["col_idx"][pat["row_ptr"][row]:pat["row_ptr"][row + 1]] == col)[0]
                blk = Jp[9 * (pat["row_ptr"][row] + ks[0]):][:9].reshape(3, 3) if len(ks) else np.zeros((3, 3))
                tol = 1e-5 * scale[var] + 1e-14 * rmax / hh
                assert np.max(np.abs(blk[:, var] - fdr[3 * row:3 * row + 3])) <= tol, (row, col, var)


def test_dispersion_conserves_and_flows_sum_to_the_residual():
    c = _case_nowells(5, 4, 3, seed=5)
    pv, mean = mixed_state(c, seed=5)
    prob = orc.problem_from_case(c)
    dsp = orc.dispersion(cell_dispersion=np.full(c.N, 1e3))
    dif = _diff()
    nv = np.random.default_rng(5).uniform(1e-7, 1e-5, c.N)
    _, r, _ = _disp_lin(prob, pv, mean, nv, dsp, dif)
    R = r.reshape(-1, 3)
    assert np.max(np.abs(R.sum(axis=0))) <= 1e-12 * np.abs(R).sum(axis=0).max()
    fl, _ = orc.flows(prob, pv, mean, diffusion=dif, dispersion=dsp, nv=nv)
    nbp = orc.pattern(prob)["nb_ptr"]
    acc = np.zeros(3 * c.N)
    for I in range(c.N):
        a = np.zeros(3)
        for k in range(nbp[I], nbp[I + 1]):
            a = a + fl[k]
        acc[3 * I:3 * I + 3] = a
    assert np.array_equal(acc.view(np.uint64), r.view(np.uint64))


def test_face_coefficients_of_cells_without_porosity_or_dispersion_are_zero():
    """A face between two cells without a dispersion coefficient has no
    dispersivity (0, not the harmonic mean's 0/0); zero coefficients
    everywhere give the residual of the model without dispersion, and
    negative or non-finite coefficients are rejected."""
    c = _case_nowells(4, 3, 2, seed=6)
    pv, mean = mixed_state(c, seed=6)
    prob = orc.problem_from_case(c)
    J0, r0 = _no_acc(prob, pv, mean, None)
    E = np.zeros(c.N)
    J, r, nv = _disp_lin(prob, pv, mean, np.full(c.N, 1e-5), orc.dispersion(cell_dispersion=E))
    Jn, rn = orc.linearize(prob, pv, mean, orc.storage(prob, pv, mean), 1e300)
    assert np.all(np.isfinite(r)) and np.all(np.isfinite(J))
    assert np.array_equal(r, rn) and np.array_equal(J, Jn)
    # half the cells without a coefficient: every face touching two of them is 0
    E[::2] = 0.0
    E[1::2] = 7.0
    J, r, nv = _disp_lin(prob, pv, mean, np.full(c.N, 1e-5), orc.dispersion(cell_dispersion=E))
    assert np.all(np.isfinite(r)) and np.all(np.isfinite(J))
    for bad in (-1.0, np.nan, np.inf):
        Eb = np.full(c.N, 3.0)
        Eb[5] = bad
        with pytest.raises(AssertionError):
            _disp_lin(prob, pv, mean, np.zeros(c.N), orc.dispersion(cell_dispersion=Eb))
    fd = np.ones(3 * c.N)
    fd[4] = -2.0
    with pytest.raises(AssertionError):
        _no_acc(prob, pv, mean, _diff(face_diffusivity=fd))
