"""Pins of the lens oracle (two-phase immiscible element path, oracle/lens.cpp)
against closed forms, invariants and finite differences (CPU only)."""
import numpy as np
import pytest

import oracle as orc

P0 = orc.lens_problem()


def small(nx=12, ny=8, **kw):
    return orc.lens_problem(nx=nx, ny=ny, **kw)


def vg(swe, alpha, n):
    m = 1.0 - 1.0 / n
    pc = (swe ** (-1.0 / m) - 1.0) ** (1.0 / n) / alpha
    krw = np.sqrt(swe) * (1.0 - (1.0 - swe ** (1.0 / m)) ** m) ** 2
    krn = (1.0 - swe) ** (1.0 / 3.0) * (1.0 - swe ** (1.0 / m)) ** (2.0 * m)
    return pc, krw, krn


def test_default_problem_matches_library_defaults():
    om = pytest.importorskip("omodels_amd")
    try:
        lp = om.lens_problem()
    except OSError:
        pytest.skip("library not built")
    for name, _ in orc.OlProblem._fields_:
        assert getattr(lp, name) == getattr(P0, name), name


def test_pattern_brute_force():
    p = small(7, 5)
    rp, ci, dg = orc.lens_pattern(p)
    nx, ny = 7, 5
    for I in range(nx * ny):
        i, j = I % nx, I // nx
        cols = sorted([I] + [I + d for d, ok in ((-1, i > 0), (1, i < nx - 1), (-nx, j > 0), (nx, j < ny - 1)) if ok])
        assert list(ci[rp[I]:rp[I + 1]]) == cols
        assert ci[dg[I]] == I


@pytest.mark.parametrize("lens", [0, 1])
@pytest.mark.parametrize("swe", [0.03, 0.2, 0.5, 0.8, 0.97])
def test_van_genuchten_closed_forms(lens, swe):
    swr = P0.swr_lens if lens else P0.swr_outer
    alpha = P0.alpha_lens if lens else P0.alpha_outer
    n = P0.n_lens if lens else P0.n_outer
    sw = swr + swe * (1.0 - swr)
    got = orc.lens_material(P0, lens, sw)
    pc, krw, krn = vg(swe, alpha, n)
    assert got["pc"] == pytest.approx(pc, rel=1e-13)
    assert got["krw"] == pytest.approx(krw, rel=1e-13)
    assert got["krn"] == pytest.approx(krn, rel=1e-13)
    # derivatives w.r.t. Sw against central differences of the closed forms
    h = 1e-7
    for key, f in (("dpc", 0), ("dkrw", 1), ("dkrn", 2)):
        a = vg(swe + h / (1 - swr), alpha, n)[f]
        b = vg(swe - h / (1 - swr), alpha, n)[f]
        assert got[key] == pytest.approx((a - b) / (2 * h), rel=2e-6)


@pytest.mark.parametrize("lens", [0, 1])
def test_regularization_is_continuous_and_monotone(lens):
    swr = P0.swr_lens if lens else P0.swr_outer

    def pc(swe):
        return orc.lens_material(P0, lens, swr + swe * (1.0 - swr))

    for s0 in (0.01, 0.99):
        lo, hi = pc(s0 - 1e-10), pc(s0 + 1e-10)
        assert lo["pc"] == pytest.approx(hi["pc"], rel=1e-7)
        assert lo["dpc"] == pytest.approx(hi["dpc"], rel=1e-6)
    # the Hermite piece reaches zero at Swe = 1 and continues linearly beyond
    assert abs(pc(1.0 - 1e-12)["pc"]) < 1e-6 * pc(0.99)["pc"]
    # dpc is d/dSw; the line is in Swe, dSwe/dSw = 1 / (1 - Swr)
    assert pc(1.05)["pc"] == pytest.approx(pc(1.0)["dpc"] * (1.0 - swr) * 0.05, rel=1e-9)
    vals = [pc(s)["pc"] for s in np.linspace(-0.05, 1.05, 221)]
    assert all(a > b for a, b in zip(vals, vals[1:]))


def test_relperm_limits():
    for lens in (0, 1):
        swr = P0.swr_lens if lens else P0.swr_outer
        dry = orc.lens_material(P0, lens, swr - 0.01)
        wet = orc.lens_material(P0, lens, 1.01)
        assert (dry["krw"], dry["krn"], dry["dkrw"], dry["dkrn"]) == (0.0, 1.0, 0.0, 0.0)
        assert (wet["krw"], wet["krn"], wet["dkrw"], wet["dkrn"]) == (1.0, 0.0, 0.0, 0.0)


def test_initial_state_is_hydrostatic_away_from_driven_boundaries():
    p = P0
    pv = orc.lens_initial(p)
    st = orc.lens_storage(p, pv)
    r, _ = orc.lens_linearize(p, pv, st, 250.0)
    r = r.reshape(-1, 2)
    nx, ny = p.nx, p.ny
    dx = p.lx / nx
    driven = np.zeros(nx * ny, bool)
    driven[0::nx] = True                       # left column: pressure scaled by 1 + 1.5/height
    xc = (np.arange(nx) + 0.5) * dx
    lam = (p.lx - xc) / p.lx
    top = (ny - 1) * nx + np.nonzero((lam > 0.5) & (lam < 2.0 / 3.0))[0]
    driven[top] = True                         # DNAPL inlet
    assert np.abs(r[~driven]).max() < 1e-9
    assert np.abs(r[driven]).min(axis=1).max() >= 0.0
    assert np.all(np.abs(r[driven]).max(axis=1) > 1e-3)


def _state_after(p, t_end):
    pv = orc.lens_initial(p)
    rc, pv, th, dth, log = orc.lens_simulate(p, pv, t_end, 250.0, orc.params(newton_tol=1e-8))
    assert rc == 0
    return pv, dth[-1]


def test_jacobian_matches_central_differences():
    p = small()
    pv, dt = _state_after(p, 2000.0)
    st1 = orc.lens_storage(p, pv) * 0.999 + 1e-3
    r0, J = orc.lens_linearize(p, pv, st1, dt)
    rp, ci, dg = orc.lens_pattern(p)
    n = p.nx * p.ny
    dense = np.zeros((2 * n, 2 * n))
    for I in range(n):
        for k in range(rp[I], rp[I + 1]):
            dense[2 * I:2 * I + 2, 2 * ci[k]:2 * ci[k] + 2] = J[4 * k:4 * k + 4].reshape(2, 2)
    checked_kink = 0
    for c in range(2 * n):
        h = 1e-3 if c % 2 == 0 else 1e-7
        if c % 2 == 1 and 1.0 - 1e-3 < pv[c] < 1.0:
            continue   # just below Swe = 1 the kr slopes are unbounded: no difference quotient resolves them
        a = pv.copy(); a[c] += h
        if c % 2 == 1 and pv[c] >= 1.0:
            # water-saturated cell: the material law changes branch at Swe = 1 and the
            # Van Genuchten kr have unbounded slopes below it, so the AD derivative is the
            # one-sided derivative of the Swe >= 1 branch: compare a forward difference
            fd = (orc.lens_linearize(p, a, st1, dt)[0] - r0) / h
            checked_kink += 1
        else:
            b = pv.copy(); b[c] -= h
            fd = (orc.lens_linearize(p, a, st1, dt)[0] - orc.lens_linearize(p, b, st1, dt)[0]) / (2 * h)
        scale = np.abs(dense[:, c]).max() + 1e-30
        assert np.abs(fd - dense[:, c]).max() <= 2e-5 * scale, c
    assert 0 < checked_kink < n


def test_column_sums_vanish_without_storage():
    """Each interior face flux enters its two cells with opposite signs, so the
    columns of J (times V) sum to zero for cells without boundary faces when the
    storage derivative is negligible (dt -> infinity)."""
    p = small()
    pv, _ = _state_after(p, 2000.0)
    st1 = orc.lens_storage(p, pv)
    _, J = orc.lens_linearize(p, pv, st1, 1e30)
    rp, ci, dg = orc.lens_pattern(p)
    n, nx, ny = p.nx * p.ny, p.nx, p.ny
    colsum = np.zeros((n, 2, 2))
    colabs = np.zeros((n, 2, 2))
    for I in range(n):
        for k in range(rp[I], rp[I + 1]):
            colsum[ci[k]] += J[4 * k:4 * k + 4].reshape(2, 2)
            colabs[ci[k]] += np.abs(J[4 * k:4 * k + 4].reshape(2, 2))
    for I in range(n):
        i, j = I % nx, I // nx
        if 0 < i < nx - 1 and 0 < j < ny - 1:
            assert np.all(np.abs(colsum[I]) <= 1e-12 * colabs[I].max())


def test_dnapl_mass_equals_inflow():
    p = P0
    pv, _ = _state_after(p, 3000.0)
    sn = 1.0 - pv[1::2]
    v = (p.lx / p.nx) * (p.ly / p.ny)
    mass = np.sum(sn * p.porosity * p.rho_n * v)
    xc = (np.arange(p.nx) + 0.5) * p.lx / p.nx
    lam = (p.lx - xc) / p.lx
    width = np.count_nonzero((lam > 0.5) & (lam < 2.0 / 3.0)) * p.lx / p.nx
    assert mass == pytest.approx(p.inflow_rate * width * 3000.0, rel=1e-8)


def test_reference_run_reaches_end_time():
    """30 000 s from dt0 = 250 s with the FP32 linear solve of the reference binary."""
    p = P0
    pv = orc.lens_initial(p)
    rc, pv, th, dth, log = orc.lens_simulate(p, pv, 30e3, 250.0, orc.params(newton_tol=1e-8, solver_float=1))
    assert rc == 0
    assert th[-1] == pytest.approx(30e3, rel=1e-12)
    assert 15 <= len(th) <= 60
    assert np.all(pv[1::2] <= 1.0 + 1e-9) and pv[1::2].min() > 0.0
