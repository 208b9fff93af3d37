"""Shared inputs of the parity tests: black-oil cases from the library's
deterministic generator and states that exercise both gas meanings."""
import numpy as np

import omodels_amd as om
import oracle as orc


def case(nx, ny, nz, seed=7, **kw):
    return om.BlackOilCase(om.make_spec(nx, ny, nz, seed=seed, **kw))


def mixed_state(c, seed=3):
    """A state with free gas (meaning Sg) in some cells and undersaturated oil
    (meaning Rs) elsewhere, perturbed pressures and water saturations."""
    rng = np.random.default_rng(seed)
    pv = c.pv0.copy()
    mean = c.mean0.copy()
    N = c.N
    pv[1::3] *= 1.0 + 0.02 * rng.standard_normal(N)
    pv[0::3] = np.clip(0.2 + 0.3 * rng.random(N), 0.0, 0.9)
    gas = rng.random(N) < 0.4
    mean[gas] = 0
    pv[2::3][gas] = 0.3 * rng.random(gas.sum()) * (1.0 - pv[0::3][gas])
    # undersaturated cells: Rs below the saturated value
    pv[2::3][~gas] *= 0.9 + 0.1 * rng.random((~gas).sum())
    return pv, mean


def rel_err(a, b, scale=None):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    s = np.max(np.abs(b)) if scale is None else scale
    return float(np.max(np.abs(a - b)) / max(s, 1e-300))


def block_rel_err(Ja, Jb):
    """max over blocks of |A - B| / max|B| of that block row (3x3 blocks)."""
    A = Ja.reshape(-1, 9)
    B = Jb.reshape(-1, 9)
    s = np.maximum(np.max(np.abs(B), axis=1), 1e-300)
    return float(np.max(np.max(np.abs(A - B), axis=1) / s))


def own_trans(prob):
    """[3][N] transmissibilities of the structured grid at the face's index
    (axis*N + I for face I|I+e_axis), from the oracle's pattern"""
    pat = orc.pattern(prob)
    N = prob.nx * prob.ny * prob.nz
    t = np.zeros(3 * N)
    for I in range(N):
        for k in range(pat["nb_ptr"][I], pat["nb_ptr"][I + 1]):
            if pat["nb_cell"][k] > I:
                t[(pat["nb_dir"][k] // 2) * N + I] = pat["nb_trans"][k]
    return t


def own_depth(c):
    s = c.spec
    return s.dz * (np.arange(c.N) // (s.nx * s.ny) + 0.5)


def random_discretization(c, seed, thp_fraction=0.3, thp_max=4e5):
    """Transmissibilities, depths, volumes and threshold pressures of a
    distorted grid: the grid's own values perturbed, some volumes equal (ties
    of the upwind decision), some faces closed, thresholds on a fraction of
    the faces."""
    rng = np.random.default_rng(seed)
    N = c.N
    V = c.spec.dx * c.spec.dy * c.spec.dz
    t = own_trans(orc.problem_from_case(c)) * (0.5 + rng.random(3 * N))
    t[rng.random(3 * N) < 0.05] = 0.0
    z = own_depth(c) + 0.3 * c.spec.dz * rng.standard_normal(N)
    vol = V * (0.5 + rng.random(N))
    vol[::7] = V
    thp = np.where(rng.random(3 * N) < thp_fraction, thp_max * rng.random(3 * N), 0.0)
    return t, z, vol, thp
