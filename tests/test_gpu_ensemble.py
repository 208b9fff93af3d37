"""Batching and serving forms: many realizations of different sizes in the
same launches, admitted by HBM budget as they arrive.

Bars: each member's converged state matches the oracle run of that member
alone (per-step pressures within 1e-8 relative at a tight Newton tolerance);
a member's result is bit-identical alone and in any mix (its arithmetic,
including the order of its reductions, does not depend on the other
members); requests arriving while others run are admitted at the next step
boundary and overlap on the device; the HBM budget bounds the batch."""
import numpy as np
import pytest

import omodels_amd as om
import oracle as orc
from cases import case

pytestmark = pytest.mark.gpu

FIXED = dict(max_dt=86400.0, min_dt=86400.0)


def _members():
    return [case(16, 12, 5, seed=31, wells=2), case(20, 10, 4, seed=32, wells=0, perm_sigma=2.0),
            case(9, 9, 6, seed=33, wells=2, perm_sigma=0.5), case(24, 16, 3, seed=34, wells=1)]


def _serve(members, params, arrivals=None, budget=64e9, max_batch=256, n_steps=3):
    srv = om.Server(params, budget, max_batch)
    ids = [srv.submit(c, n_steps, 86400.0, 0.0 if arrivals is None else arrivals[k]) for k, c in enumerate(members)]
    srv.run()
    return srv, [srv.result(i) for i in ids]


def test_ensemble_members_match_the_oracle():
    ms = _members()
    srv, res = _serve(ms, om.default_params(newton_tol=1e-10, **FIXED))
    for c, r in zip(ms, res):
        assert r["status"] == 0
        rco, pvo, mo, histo, dtso, logo, tim = orc.simulate(orc.problem_from_case(c), c.pv0, c.mean0, 3, 86400.0,
                                                            orc.params(newton_tol=1e-10, **FIXED), nthreads=4)
        assert rco == 0
        p, po = r["pv"][1::3], pvo[1::3]
        assert np.max(np.abs(p - po)) <= 1e-8 * np.max(np.abs(po))
        assert np.max(np.abs(r["pv"][0::3] - pvo[0::3])) <= 1e-8
        assert np.array_equal(r["mean"], mo)


@pytest.mark.parametrize("schedule", [0, 2, 3])
@pytest.mark.parametrize("precision", [0, 1])
def test_member_result_is_identical_alone_and_in_the_mix(precision, schedule):
    """schedule 0 (batched sweeps for ensembles), 2 / 3 force the streaming /
    batched sweeps: the arithmetic is the same in every case."""
    ms = _members()
    prm = om.default_params(precision=precision, ilu_schedule=schedule)
    srv, mixed = _serve(ms, prm)
    for k, c in enumerate(ms):
        srv1, alone = _serve([c], prm)
        assert alone[0]["status"] == 0 and mixed[k]["status"] == 0
        assert np.array_equal(alone[0]["pv"].view(np.uint64), mixed[k]["pv"].view(np.uint64))
        assert np.array_equal(alone[0]["mean"], mixed[k]["mean"])
        assert np.array_equal(alone[0]["dts"], mixed[k]["dts"])


def test_mixed_precision_ensemble_matches_fp64_oracle():
    ms = _members()
    kw = dict(newton_tol=1e-7, **FIXED)
    srv, res = _serve(ms, om.default_params(precision=1, **kw))
    for c, r in zip(ms, res):
        rco, pvo, mo, histo, dtso, logo, tim = orc.simulate(orc.problem_from_case(c), c.pv0, c.mean0, 3, 86400.0,
                                                            orc.params(**kw), nthreads=4)
        assert r["status"] == 0 and rco == 0
        assert np.max(np.abs(r["pv"][1::3] - pvo[1::3])) <= 1e-6 * np.max(np.abs(pvo[1::3]))


def test_continuous_batching_admits_arrivals_while_others_run():
    ms = _members()
    prm = om.default_params()
    # the first two arrive at once; the others arrive while those are running
    srv, res = _serve(ms, prm, arrivals=[0.0, 0.0, 0.05, 0.1], n_steps=6)
    for k, r in enumerate(res):
        assert r["status"] == 0
        assert r["admit"] >= r["arrival"]
        assert r["finish"] > r["admit"]
    s = srv.summary()
    assert s["max_running"] >= 2
    # a late request is admitted before an earlier one finished: their work overlapped
    assert min(res[2]["admit"], res[3]["admit"]) < max(res[0]["finish"], res[1]["finish"])
    for k, c in enumerate(ms):
        srv1, alone = _serve([c], prm, n_steps=6)
        assert np.array_equal(alone[0]["pv"].view(np.uint64), res[k]["pv"].view(np.uint64))


def test_hbm_budget_limits_the_batch():
    ms = _members()
    prm = om.default_params()
    b = sorted(om_bytes(c) for c in ms)
    budget = 1.01 * max(b[-1], b[0] + b[1])     # any two fit, no three do
    assert b[0] + b[1] + b[2] > budget
    srv, res = _serve(ms, prm, budget=budget)
    assert all(r["status"] == 0 for r in res)
    assert srv.summary()["max_running"] <= 2


def om_bytes(c):
    N = c.N
    # same estimate as the library's admission control for FP64
    nnzb = 7 * N
    return nnzb * 9 * 8 * 2 + N * 9 * 8 + N * 3 * 8 * (8 + 5) + N * 3 * 4 * 8 + N * (4 * 8 + 6 * 4 + 3)
