"""The committed golden vectors (tests/golden, written by make_golden.py from
the oracle) reproduced by the oracle (CPU) and by the HIP library (GPU)."""
import numpy as np
import pytest

import omodels_amd as om
import oracle as orc
from golden_case import GoldenCase

NAMES = ["blackoil_4x3x2.npz", "blackoil_9x7x4.npz"]


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_golden(name):
    g = GoldenCase(name)
    z = g.z
    prob = orc.problem_from_case(g)
    pat = orc.pattern(prob)
    for k in ("row_ptr", "col_idx", "diag_off", "nb_ptr", "nb_cell", "nb_off"):
        assert np.array_equal(pat[k], z[k]), k
    J, r = orc.linearize(prob, z["state_pv"], z["state_mean"], z["storage1"], float(z["dt"][0]))
    assert np.array_equal(_bits(J), _bits(z["jac"])) and np.array_equal(_bits(r), _bits(z["res"]))
    rc, lu = orc.ilu0(g.N, 3, pat["row_ptr"], pat["col_idx"], J)
    assert np.array_equal(_bits(lu), _bits(z["ilu"]))
    rc, pvf, mf, hist, dts, log, tim = orc.simulate(prob, g.pv0, g.mean0, 3, 86400.0,
                                                    orc.params(newton_tol=1e-10, max_dt=86400.0, min_dt=86400.0))
    assert rc == 0
    assert np.array_equal(_bits(hist), _bits(z["hist"]))
    rc, pvf, mf, hist, dts, log, tim = orc.simulate(prob, g.pv0, g.mean0, 5, 86400.0, orc.params())
    assert rc == 0
    assert np.array_equal(_bits(hist), _bits(z["hist_ref"])) and np.array_equal(_bits(dts), _bits(z["hist_ref_dt"]))
    assert [tuple(l[:6]) for l in log] == [tuple(l) for l in z["log_ref"]]


@pytest.mark.parametrize("name", NAMES)
def test_golden_records_the_chosen_forms(name):
    forms = list(GoldenCase(name).z["forms"])
    assert any(f.startswith("scalar product: exact sum") for f in forms)
    assert any(f.startswith("LiveOilPvt undersaturated") for f in forms)


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_gpu_reproduces_golden(name):
    g = GoldenCase(name)
    z = g.z
    m = om.Model(g)
    pat = m.pattern()
    for k in ("row_ptr", "col_idx", "diag_off", "nb_ptr", "nb_cell", "nb_off"):
        assert np.array_equal(pat[k], z[k]), k
    assert np.array_equal(_bits(pat["nb_trans"]), _bits(z["nb_trans"]))
    m.set_solution(z["state_pv"], z["state_mean"])
    m.set_storage1(np.ascontiguousarray(z["storage1"]))
    J, r = m.linearize(float(z["dt"][0]), iteration_index=1)
    assert np.array_equal(_bits(J), _bits(z["jac"])) and np.array_equal(_bits(r), _bits(z["res"]))
    s = om.Solver(m, om.default_params())
    s.prepare(np.ascontiguousarray(z["jac"]), np.ascontiguousarray(z["res"]))
    assert np.array_equal(_bits(s.ilu()), _bits(z["ilu"]))
    m2 = om.Model(g)
    m2.set_solution(g.pv0, g.mean0)
    rc, hist, dts, log, st = m2.simulate(om.default_params(newton_tol=1e-10, max_dt=86400.0, min_dt=86400.0), 3,
                                         86400.0)
    assert rc == 0
    p, po = hist[:, 1::3], z["hist"][:, 1::3]
    assert np.max(np.abs(p - po)) <= 1e-8 * np.max(np.abs(po))
    assert np.max(np.abs(hist[:, 0::3] - z["hist"][:, 0::3])) <= 1e-8
    assert [tuple(l[:6]) for l in log] == [tuple(l) for l in z["log"]]
    assert np.array_equal(_bits(hist), _bits(z["hist"]))
    # the reference's own settings: step sizes, iteration counts and states bit for bit
    m3 = om.Model(g)
    m3.set_solution(g.pv0, g.mean0)
    rc, hist, dts, log, st = m3.simulate(om.default_params(), 5, 86400.0)
    assert rc == 0
    assert np.array_equal(_bits(dts), _bits(z["hist_ref_dt"]))
    assert [tuple(l[:6]) for l in log] == [tuple(l) for l in z["log_ref"]]
    assert np.array_equal(_bits([l[6] for l in log]), _bits(z["log_ref_error"]))
    assert np.array_equal(_bits(hist), _bits(z["hist_ref"]))
