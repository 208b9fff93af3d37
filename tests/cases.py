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


class RateWell:
    """An auxiliary module for the tests (BaseAuxiliaryModule's interface): one
    auxiliary DOF per well whose first unknown is the well's bottom-hole
    pressure p_w; each perforated cell k takes q_k = rho WI (p_w - p_k) of
    water (kg/s, into the cell), and the well's equation holds the total at
    its target, sum_k q_k - Q = 0.  The DOF's other two unknowns are pinned
    to zero (residual x, unit diagonal).  Everything is linear in p_w and the
    cells' pressures, so the terms and their derivatives are closed forms."""

    RHO = 1000.0

    def __init__(self, N, wells):
        """wells: list of (perforated cells, well indices, target rate Q)"""
        self.N = N
        self.wells = [(np.asarray(c, np.int64), np.asarray(wi, np.float64), float(q)) for c, wi, q in wells]
        self.n_aux = len(self.wells)

    def links(self):
        rows, cols = [], []
        for w, (cells, _, _) in enumerate(self.wells):
            a = self.N + w
            rows.append(a); cols.append(a)
            for k in cells:
                rows += [int(k), a]
                cols += [a, int(k)]
        return np.array(rows, np.int32), np.array(cols, np.int32)

    def initial(self, pv):
        x = np.zeros(3 * self.n_aux)
        for w, (cells, _, _) in enumerate(self.wells):
            x[3 * w] = float(np.mean(pv[3 * cells + 1])) + 1e5
        return x

    def terms(self, pv, mean, aux, iteration=0):
        """(brow, bcol, blocks, rrow, rvals) of the module's linearize"""
        brow, bcol, blocks, rrow, rvals = [], [], [], [], []
        for w, (cells, wi, q) in enumerate(self.wells):
            a = self.N + w
            pw = aux[3 * w]
            total = 0.0
            dpw = 0.0
            for k, wik in zip(cells, wi):
                rwi = self.RHO * wik
                qk = rwi * (pw - pv[3 * k + 1])
                total += qk
                dpw += rwi
                # the cell's water equation (index 1): -q_k, a source
                rrow.append(int(k)); rvals.append([0.0, -qk, 0.0])
                d = np.zeros(9); d[3 * 1 + 1] = rwi          # d(-q_k)/dp_k
                brow.append(int(k)); bcol.append(int(k)); blocks.append(d)
                o = np.zeros(9); o[3 * 1 + 0] = -rwi          # d(-q_k)/dp_w
                brow.append(int(k)); bcol.append(a); blocks.append(o)
                c = np.zeros(9); c[0 * 3 + 1] = -rwi          # d(sum q)/dp_k
                brow.append(a); bcol.append(int(k)); blocks.append(c)
            rrow.append(a); rvals.append([total - q, aux[3 * w + 1], aux[3 * w + 2]])
            dd = np.zeros(9); dd[0] = dpw; dd[4] = 1.0; dd[8] = 1.0
            brow.append(a); bcol.append(a); blocks.append(dd)
        return (np.array(brow, np.int32), np.array(bcol, np.int32), np.array(blocks).reshape(-1, 9),
                np.array(rrow, np.int32), np.array(rvals).reshape(-1, 3))
