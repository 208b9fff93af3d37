"""CPU tests of rock compressibility and compaction in the oracle
(blackoilintensivequantities.hh:447-480; the face's mean transmissibility
multiplier, blackoillocalresidualtpfa.hh:372-381), pinned by closed forms of
the porosity and its pressure derivative, the accumulation term, the face
rates with a constant and a pressure-dependent multiplier, the reference's
defaults (no bit changes), and central differences of the Jacobian."""
import numpy as np
import pytest

import omodels_amd as om
import oracle as orc
from cases import case, mixed_state

MD = 9.869233e-16


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def rock_table(pmin=100e5, pmax=300e5, n=6, pv_slope=0.02, tm_slope=0.15):
    p = np.linspace(pmin, pmax, n)
    u = (p - pmin) / (pmax - pmin)
    return np.concatenate([p, 1.0 + pv_slope * u ** 2, 0.9 + tm_slope * u ** 1.5])


def _interp(tab, x):
    """piecewise-linear [2][n] table (x, y), linear beyond the ends: value and slope"""
    n = len(tab) // 2
    p, y = tab[:n], tab[n:]
    k = min(max(np.searchsorted(p, x, side="right") - 1, 0), n - 2)
    s = (y[k + 1] - y[k]) / (p[k + 1] - p[k])
    return y[k] + s * (x - p[k]), s


def test_porosity_closed_form_and_derivative():
    c = case(2, 2, 1, seed=3)
    N = c.N
    comp = np.full(N, 4e-10)
    pref = np.full(N, 150e5)
    tab = rock_table()
    tab3 = tab.reshape(3, -1)
    prob = orc.with_rock(orc.problem_from_case(c), comp, pref, tab)
    for cell, p in ((0, 180e5), (1, 240e5), (3, 90e5)):     # the last below the table: linear extrapolation
        q = orc.intensive(prob, np.array([0.3, p, 50.0]), 1, cell=cell)
        x = comp[cell] * (p - pref[cell])
        pm, dpm = _interp(np.concatenate([tab3[0], tab3[1]]), p)
        phi0 = c.poro[cell]
        want = phi0 * (1 + x + 0.5 * x * x) * pm
        dwant = phi0 * (comp[cell] * (1 + x)) * pm + phi0 * (1 + x + 0.5 * x * x) * dpm
        assert q["phi"][0] == pytest.approx(want, rel=1e-14)
        assert q["phi"][2] == pytest.approx(dwant, rel=1e-12)      # d/dp_o
        assert q["phi"][1] == 0.0 and q["phi"][3] == 0.0


def test_accumulation_scales_with_the_rock_porosity():
    c = case(3, 2, 2, seed=5)
    N = c.N
    pv, mean = mixed_state(c, seed=5)
    comp = np.random.default_rng(5).uniform(0.0, 1e-9, N)
    comp[::3] = 0.0                                 # cells without compressibility
    plain = orc.problem_from_case(c)
    prob = orc.with_rock(orc.problem_from_case(c), comp, None)
    st0, st = orc.storage(plain, pv, mean), orc.storage(prob, pv, mean)
    x = comp * (pv[1::3] - 1e5)                     # default reference pressure
    mult = 1 + x + 0.5 * x * x
    assert np.allclose(st, st0 * np.repeat(mult, 3), rtol=1e-14, atol=0)
    assert np.array_equal(_bits(st[0::9]), _bits(st0[0::9]))      # c = 0: untouched


def test_reference_defaults_change_no_bit():
    """compressibility 0 and multipliers 1 are the reference problem's
    defaults: the rock terms then change no bit of J, r or the rates"""
    c = case(5, 4, 3, seed=7)
    pv, mean = mixed_state(c, seed=7)
    plain = orc.problem_from_case(c)
    st1 = orc.storage(plain, *mixed_state(c, seed=8))
    J0, r0 = orc.linearize(plain, pv, mean, st1, 43200.0)
    f0, d0 = orc.flows(plain, pv, mean)
    ones = np.concatenate([np.array([100e5, 400e5]), np.ones(4)])
    prob = orc.with_rock(orc.problem_from_case(c), np.zeros(c.N), np.full(c.N, 2e5), ones)
    J1, r1 = orc.linearize(prob, pv, mean, st1, 43200.0)
    f1, d1 = orc.flows(prob, pv, mean)
    for a, b in ((J0, J1), (r0, r1), (f0, f1), (d0, d1)):
        assert np.array_equal(_bits(a), _bits(b))


def test_face_rates_carry_the_mean_transmissibility_multiplier():
    c = case(5, 4, 2, seed=9)
    N = c.N
    pv, mean = mixed_state(c, seed=9)
    plain = orc.problem_from_case(c)
    _, fr0 = orc.flows(plain, pv, mean)
    # a constant multiplier 2: every rate doubles, exactly
    two = np.concatenate([np.array([100e5, 400e5]), np.ones(2), np.full(2, 2.0)])
    _, fr2 = orc.flows(orc.with_rock(orc.problem_from_case(c), None, None, two), pv, mean)
    assert np.array_equal(_bits(fr2), _bits(2.0 * fr0))
    # pressure-dependent: the mean of the two cells' multipliers
    tab = rock_table()
    tab3 = tab.reshape(3, -1)
    _, fr = orc.flows(orc.with_rock(orc.problem_from_case(c), None, None, tab), pv, mean)
    pat = orc.pattern(plain)
    tmt = np.concatenate([tab3[0], tab3[2]])
    for I in range(N):
        for k in range(pat["nb_ptr"][I], pat["nb_ptr"][I + 1]):
            J = pat["nb_cell"][k]
            m = 0.5 * (_interp(tmt, pv[3 * I + 1])[0] + _interp(tmt, pv[3 * J + 1])[0])
            assert np.allclose(fr[k], fr0[k] * m, rtol=1e-13, atol=0)


@pytest.mark.parametrize("seed", [1, 2])
def test_jacobian_with_rock_matches_central_differences(seed):
    c = case(3, 3, 2, seed=seed)
    N = c.N
    pv, mean = mixed_state(c, seed=seed)
    rng = np.random.default_rng(seed)
    prob = orc.with_rock(orc.problem_from_case(c), rng.uniform(1e-10, 1e-9, N), rng.uniform(1e5, 200e5, N),
                         rock_table())
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


@pytest.mark.parametrize("bad", ["comp", "table_x", "table_y"])
def test_invalid_rock_is_rejected(bad):
    c = case(2, 2, 1)
    comp, tab = np.zeros(c.N), rock_table()
    if bad == "comp":
        comp[1] = -1e-9
    elif bad == "table_x":
        tab[2] = tab[1]
    else:
        tab[8] = 0.0
    prob = orc.with_rock(orc.problem_from_case(c), comp, None, tab)
    st = np.zeros(3 * c.N)
    assert orc.lib().ob_storage(orc.C.byref(prob), orc.P(c.pv0), orc.P(c.mean0, orc._u8p), orc.P(st)) == -1
