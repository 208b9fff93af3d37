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


@pytest.mark.parametrize("ilu_order,graph", [(0, 0), (1, 0), (1, 1)])
def test_ensemble_members_match_the_oracle(ilu_order, graph):
    """At the reference's settings (NewtonTolerance 1e-6, adaptive steps from
    one day) every member follows the oracle run of that member alone: step
    sizes, iteration counts and final state bit for bit (scalar products are
    exact sums), with the natural-order and the two-colour ILU(0), and with
    the batched BiCGStab iterations replayed from hipGraphs."""
    ms = _members()
    srv, res = _serve(ms, om.default_params(ilu_order=ilu_order, graph=graph))
    for c, r in zip(ms, res):
        assert r["status"] == 0
        rco, pvo, mo, histo, dtso, logo, tim = orc.simulate(orc.problem_from_case(c), c.pv0, c.mean0, 3, 86400.0,
                                                            orc.params(ilu_order=ilu_order), nthreads=4)
        assert rco == 0
        p, po = r["pv"][1::3], pvo[1::3]
        assert np.max(np.abs(p - po)) <= 1e-8 * np.max(np.abs(po))
        assert np.max(np.abs(r["pv"][0::3] - pvo[0::3])) <= 1e-8
        assert np.array_equal(r["mean"], mo)
        assert np.array_equal(r["dts"], dtso)
        assert np.array_equal(r["pv"].view(np.uint64), pvo.view(np.uint64))


@pytest.mark.parametrize("schedule,ilu_order,graph", [(0, 0, 0), (2, 0, 0), (3, 0, 0), (0, 1, 0), (0, 1, 1)])
@pytest.mark.parametrize("precision", [0, 1])
def test_member_result_is_identical_alone_and_in_the_mix(precision, schedule, ilu_order, graph):
    """schedule 0 (batched sweeps for ensembles), 2 / 3 force the streaming /
    batched sweeps, ilu_order 1 the two-colour ILU(0): the arithmetic is the
    same in every case, also with the BiCGStab iterations replayed from
    hipGraphs."""
    ms = _members()
    prm = om.default_params(precision=precision, ilu_schedule=schedule, ilu_order=ilu_order, graph=graph)
    srv, mixed = _serve(ms, prm)
    for k, c in enumerate(ms):
        srv1, alone = _serve([c], prm)
        assert alone[0]["status"] == 0 and mixed[k]["status"] == 0
        assert np.array_equal(alone[0]["pv"].view(np.uint64), mixed[k]["pv"].view(np.uint64))
        assert np.array_equal(alone[0]["mean"], mixed[k]["mean"])
        assert np.array_equal(alone[0]["dts"], mixed[k]["dts"])


@pytest.mark.parametrize("ilu_order", [0, 1])
def test_mixed_precision_ensemble_matches_fp64_oracle(ilu_order):
    ms = _members()
    kw = dict(newton_tol=1e-7, **FIXED)
    srv, res = _serve(ms, om.default_params(precision=1, ilu_order=ilu_order, **kw))
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


def test_members_step_on_their_own():
    """The batch runs rounds of one Newton iteration per running member: a
    member needing fewer iterations for its steps finishes (and leaves) before
    the slowest one instead of waiting at every step boundary, and the number
    of rounds is the slowest member's number of linearizations."""
    ms = _members()
    srv, res = _serve(ms, om.default_params(), n_steps=4)
    assert all(r["status"] == 0 for r in res)
    lins = [r["stats"].linearizations for r in res]
    assert min(lins) < max(lins)
    assert srv.summary()["batch_rounds"] == max(lins)
    slow = int(np.argmax(lins))
    for k, r in enumerate(res):
        if lins[k] < lins[slow]:
            assert r["finish"] < res[slow]["finish"], (k, lins)


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


def test_exhausted_ilu_wait_fails_the_solve_and_the_next_runs_clean():
    """A dependency wait of the tiled ILU that ran into its bound (forced here
    by setting the flag from the host, so every wait gives up and every kernel
    ends normally) fails the batched solve of every member as a
    NumericalProblem: each member halves its step and goes on.  The next
    request on the same server runs clean: identical to a fresh server."""
    ms = _members()[:2]
    prm = om.default_params()
    srv = om.Server(prm, 64e9, 64)
    ids = [srv.submit(c, 2, 86400.0, 0.0) for c in ms]
    om.check(om.lib().omodels_server_inject_ilu_wait_error(srv.h), "inject")
    srv.run()
    res = [srv.result(i) for i in ids]
    for r in res:
        assert r["status"] == 0
        assert r["stats"].failed_attempts >= 1
        assert r["dts"][0] == 43200.0
    late = srv.submit(ms[0], 2, 86400.0, 0.0)
    srv.run()
    again = srv.result(late)
    srv1, fresh = _serve([ms[0]], prm, n_steps=2)
    assert again["status"] == 0 and again["stats"].failed_attempts == 0
    assert np.array_equal(again["pv"].view(np.uint64), fresh[0]["pv"].view(np.uint64))


def test_serving_members_of_8k_to_60k_cells():
    """Members of the serving workload's sizes (8k to 60k cells, uneven
    shapes, their own permeability fields and wells), FP32-inner with the
    two-colour ILU(0), two one-day steps: each result is bit-identical alone
    and in the mix, and within 1e-6 relative of the FP64 oracle in pressure."""
    specs = [(20, 20, 20, 41, 0.5), (40, 25, 20, 42, 1.2), (50, 40, 20, 43, 2.0), (60, 50, 20, 44, 1.0)]
    ms = [case(nx, ny, nz, seed=sd, wells=2, perm_sigma=sg) for nx, ny, nz, sd, sg in specs]
    assert min(c.N for c in ms) >= 8000 and max(c.N for c in ms) >= 60000
    kw = dict(newton_tol=1e-7, **FIXED)
    prm = om.default_params(precision=1, ilu_order=1, **kw)
    srv, mixed = _serve(ms, prm, n_steps=2)
    for k, c in enumerate(ms):
        srv1, alone = _serve([c], prm, n_steps=2)
        assert alone[0]["status"] == 0 and mixed[k]["status"] == 0
        assert np.array_equal(alone[0]["pv"].view(np.uint64), mixed[k]["pv"].view(np.uint64))
        rco, pvo, mo, histo, dtso, logo, tim = orc.simulate(orc.problem_from_case(c), c.pv0, c.mean0, 2, 86400.0,
                                                            orc.params(**kw), nthreads=8)
        assert rco == 0
        p, po = mixed[k]["pv"][1::3], pvo[1::3]
        assert np.max(np.abs(p - po)) <= 1e-6 * np.max(np.abs(po)), k
