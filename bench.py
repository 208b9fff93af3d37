#!/usr/bin/env python3
"""Benchmark of the fully-implicit black-oil hot path on one MI355X.

Workload (the headline metric's configuration): black-oil, 100x100x50 cells
(500k) of 50x50x5 m, log-normal kx (mean 100 mD, sigma 1), kz = 0.1 kx,
porosity U(0.15, 0.30), hydrostatic saturated-oil initial state at 200 bar,
corner water injector and BHP producer; FP64 TPFA assembly, block-ILU0
BiCGStab and the black-oil Newton method with the reference's parameters.

A "step" is one implicit time step of one day (the Newton loop until
convergence); value = Newton iterations per second over the timed steps.
Kernel durations come from HIP events recorded on the library's stream.

The same run also serves an ensemble request stream (the serving workload):
black-oil realizations of 8k-250k cells with their own seeded permeability
(sigma 0.5-2.0) and random well placement, Poisson arrivals, 10 time steps of
at most one day each (halved on a failed Newton solve), admitted by HBM budget into the continuous-batching server and
solved with FP32 inner / FP64 outer BiCGStab.  It reports solves/s and p50/p99
latency (completion - arrival), and the plain FP64 single-reservoir path on the
first requests of the same stream for comparison (solves/s, and the pressure
difference of the served results against it).

  python bench.py [--steps K] [--warmup W] [--gpus 1] [--config 500k|c2|c3]
                  [--precision 0|1] [--ilu-order 0|1] [--precond 0|2] [--no-cpu] [--no-serve]
                  [--serve-requests R] [--serve-rate REQ_PER_S] [--serve-plain P]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "opm-models_amd"))

HBM_PEAK_GBS = 8000.0

CONFIGS = {
    "500k": dict(nx=100, ny=100, nz=50, label="blackoil 100x100x50 (500k cells), 3x3 BCRS"),
    "c2": dict(nx=40, ny=40, nz=10, label="blackoil 40x40x10 (16k cells), 3x3 BCRS"),
    "c3": dict(nx=120, ny=120, nz=30, label="blackoil 120x120x30 (432k cells), 3x3 BCRS"),
    # 2M cells with the headline's log-normal field and corner wells; the
    # injection rate grows with the pore volume, so the reference's adaptive
    # step control (at most one day, halved on failure) sets the steps
    "c4": dict(nx=200, ny=200, nz=50, adaptive=True, label="blackoil 200x200x50 (2M cells), 3x3 BCRS"),
    # channelised field (1000 mD channels in a 1 mD background, sigma 2 within
    # each facies), five-spot with the injector in a channel (250 kg/s split
    # over its layers by kh), oil 10% undersaturated at the start, the
    # reference's adaptive step control
    "c4ch": dict(nx=200, ny=200, nz=50, spec=dict(field=1, perm_sigma=2.0, wells=1, inj_rate=250.0, rs_frac=0.9),
                 adaptive=True, label="blackoil 200x200x50 (2M cells) channelised five-spot, 3x3 BCRS"),
}


# HBM bytes per launch at the 500k workload, from two rocprofv3 --pmc passes
# (FETCH_SIZE, then WRITE_SIZE) scaled by what the same counters read for a
# device copy of 256 MiB (FETCH_SIZE x 2.0, WRITE_SIZE x 1.0); the commands and
# per-kernel figures are in profiles/blackoil_500k_hbm_traffic.md.
HBM_TRAFFIC_500K = {
    "ilu_apply": 585870950,    # forward + backward sweep
    "spmv": 290598912,
    "assemble": 489551360,
    "ilu_factor": 634073804,
}
# the same at 500k with the FP32 inner solve and the two-colour ILU
# (profiles/blackoil_500k_fp32_two_colour.md)
HBM_TRAFFIC_500K_FP32_COLORED = {
    "ilu_apply": 387115213,    # both colours' forward and backward sweeps
    "spmv": 138695680,
    "ilu_factor": 296877773,
}


def note(msg):
    """a progress line on stderr (the JSON line alone goes to stdout)"""
    print("[bench] " + msg, file=sys.stderr, flush=True)


def algorithmic_bytes(N, F, nnzb, n=3, s=8, colored=False):
    """Per-launch algorithmic HBM bytes of each kernel class; s = bytes of the
    inner solve's scalar (8, or 4 with the FP32 inner solve), colored = the
    two-colour ILU (FP64 pivots, off-diagonal blocks of the solver's matrix)."""
    ab = {
        # primaries, old storage, porosity, depth, meaning flags; one T per face;
        # residual and Jacobian (FP64) writes
        "assemble": N * (2 * n * 8 + 8 + 8 + 4) + F * 8 + N * n * 8 + nnzb * n * n * 8,
        "spmv": nnzb * (n * n * s + 4) + (N + 1) * 4 + 2 * N * n * s,
        "ilu_apply": nnzb * (n * n * s + 4) + (N + 1) * 4 + 3 * N * n * s,
        # FP64 factor in both precisions: read and write every block once
        "ilu_factor": 2 * nnzb * n * n * 8 + nnzb * 4 + (N + 1) * 4,
        # preconditioner application with the operator application in its backward sweep
        "ilu_apply_spmv": 2 * (nnzb * (n * n * s + 4) + (N + 1) * 4) + 5 * N * n * s,
    }
    if colored:
        # both colours' forward and backward sweeps: every off-diagonal block
        # once, the FP64 pivots twice, d read and v written in each direction
        ab["ilu_apply"] = (nnzb - N) * n * n * s + 2 * 8 * n * n * N + 4 * n * s * N
        # the blocks of A once (FP64), the pivots written
        ab["ilu_factor"] = nnzb * n * n * 8 + N * n * n * 8
    return ab


def amg_levels(nx, ny, nz, agg=(3, 3, 3), target=32):
    """level dimensions of the aggregation AMG (precond 2): boxes of agg cells
    until a level has at most target cells"""
    d = [(nx, ny, nz)]
    while d[-1][0] * d[-1][1] * d[-1][2] > target and len(d) < 16:
        c = tuple(-(-d[-1][q] // agg[q]) for q in range(3))
        if c == d[-1]:
            break
        d.append(c)
    return d


def amg_bytes(levels, pre=2, post=2, n=3, s=8):
    """algorithmic HBM bytes of one V-cycle and of one setup (s: bytes of the
    level matrices' and vectors' scalar; inverted diagonal blocks and the
    coarsest inverse are FP64).  Per level of N cells and NNZB blocks: a
    sweep reads the off-diagonal blocks, the inverted diagonal blocks, b and x
    and writes x (the first, from x = 0, reads only the blocks of neighbours
    swept before a cell: half of the off-diagonal ones, and no x); the
    residual reads the blocks, b and x and writes r; the restriction reads r;
    the prolongation reads and writes x.  The coarsest level reads its dense
    inverse.  Setup (FP64): every level's matrix read once (structured
    layout, 7 slots), coarse matrices and inverted diagonal blocks written."""
    bb = n * n * s
    cyc = 0
    setup = 0
    for l, (nx, ny, nz) in enumerate(levels):
        N = nx * ny * nz
        nnzb = N + 2 * ((nx - 1) * ny * nz + nx * (ny - 1) * nz + nx * ny * (nz - 1))
        vec = n * s * N
        piv = n * n * 8 * N
        bb64 = n * n * 8
        if l + 1 == len(levels):
            cyc += (n * N) ** 2 * 8 + 2 * vec
            setup += 7 * bb64 * N + (n * N) ** 2 * 8
            continue
        Nc = 1
        for q in levels[l + 1]:
            Nc *= q
        zero_sweep = ((nnzb - N) // 2) * bb + piv + 2 * vec
        sweep = (nnzb - N) * bb + piv + 3 * vec
        cyc += zero_sweep + (pre - 1 + post) * sweep
        cyc += nnzb * bb + 3 * vec                 # residual
        cyc += vec + n * s * Nc                    # restriction
        cyc += 2 * vec + n * s * Nc                # prolongation
        setup += 7 * bb64 * N + piv + 7 * bb64 * Nc
    return cyc, setup


def serve_workload(om, np, n_req, rate, n_steps, n_plain, seed=77, progress=False):
    """Poisson stream of uneven black-oil realizations through the server
    with each batched BiCGStab iteration replayed from a hipGraph, then the
    same stream without the graphs and, request by request, through the plain
    FP64 path (progress: the server prints a line per finished request)."""
    rng = np.random.default_rng(seed)
    cases = []
    for k in range(n_req):
        n = rng.uniform(8e3, 250e3)
        nz = int(rng.integers(5, 21))
        nxy = max(4, int(round(np.sqrt(n / nz))))
        spec = om.make_spec(nxy, nxy, nz, seed=10007 + k, wells=2, perm_sigma=float(rng.uniform(0.5, 2.0)))
        cases.append(om.BlackOilCase(spec))
    gaps = rng.exponential(1.0 / rate, n_req)
    arrivals = np.cumsum(gaps) - gaps[0]
    dt = 86400.0
    # the reference's time-step control with steps of at most one day: a step
    # whose Newton solve fails is retried at half the size
    budget = 64 * 2.0 ** 30

    def served(graph):
        prm = om.default_params(max_dt=dt, newton_tol=1e-7, precision=1, ilu_order=1, graph=graph,
                                verbose=1 if progress else 0)
        srv = om.Server(prm, budget, 64)
        for c, a in zip(cases, arrivals):
            srv.submit(c, n_steps, dt, float(a))
        t0 = time.time()
        srv.run()
        wall = time.time() - t0
        res = [srv.result(i) for i in range(n_req)]
        summ = srv.summary()
        srv.close()
        return wall, res, summ

    wall0, res0, _ = served(0)
    note("served %d requests without graphs in %.1f s" % (n_req, wall0))
    wall, res, summ = served(1)
    note("served %d requests with graphs in %.1f s" % (n_req, wall))
    lat = np.array([r["finish"] - r["arrival"] for r in res])
    lat0 = np.array([r["finish"] - r["arrival"] for r in res0])
    same = all(np.array_equal(a["pv"].view(np.uint64), b["pv"].view(np.uint64)) for a, b in zip(res, res0))
    out = {
        "requests": n_req, "steps_per_request": n_steps, "max_dt_days": 1.0,
        "cells_min": int(min(c.N for c in cases)), "cells_max": int(max(c.N for c in cases)),
        "cells_total": int(sum(c.N for c in cases)),
        "arrival_rate_per_s": rate, "precision": "fp32-inner/fp64-outer",
        "precond": "block-ILU0, two-colour 4x2x8 tiles", "newton_tol": 1e-7,
        "hbm_budget_GiB": 64, "max_batch": 64, "max_running": summ["max_running"],
        "batch_newton_rounds": summ["batch_rounds"], "stepping": "per member (rounds of one Newton iteration)",
        "graph": "each batched BiCGStab half-iteration replayed from a hipGraph",
        "wall_s": round(wall, 3), "solves_per_s": round(n_req / wall, 4),
        "p50_latency_s": round(float(np.percentile(lat, 50)), 3),
        "p99_latency_s": round(float(np.percentile(lat, 99)), 3),
        "without_graph": {"wall_s": round(wall0, 3), "solves_per_s": round(n_req / wall0, 4),
                          "p50_latency_s": round(float(np.percentile(lat0, 50)), 3),
                          "p99_latency_s": round(float(np.percentile(lat0, 99)), 3),
                          "results_bit_identical": bool(same)},
        "failed": int(sum(1 for r in res if r["status"] != 0)),
        "newton_iters": int(sum(r["stats"].newton_iters for r in res)),
        "linear_iters": int(sum(r["stats"].linear_iters for r in res)),
    }
    if n_plain > 0:
        # the plain FP64 drop-in path, one reservoir at a time, same requests
        # and the same time-step control
        pprm = om.default_params(max_dt=dt, newton_tol=1e-7, precision=0)
        t0 = time.time()
        finals = []
        for k in range(n_plain):
            m = om.Model(cases[k])
            m.set_solution(cases[k].pv0, cases[k].mean0)
            rc, _, dts, _, _ = m.simulate(pprm, n_steps, dt, keep_history=False)
            if k % 8 == 7:
                note("plain FP64: %d of %d requests" % (k + 1, n_plain))
            if rc != 0:
                raise SystemExit("plain FP64 run of request %d failed rc=%d" % (k, rc))
            finals.append((m.get_solution()[0], dts))
            m.close()
        pw = time.time() - t0
        # accuracy at equal times: where the two runs chose different step
        # sizes, the FP64 path is replayed with the served run's steps
        errs, replayed = [], 0
        for k in range(n_plain):
            pv, dts = finals[k]
            if not np.array_equal(dts, res[k]["dts"]):
                replayed += 1
                m = om.Model(cases[k])
                m.set_solution(cases[k].pv0, cases[k].mean0)
                for h in res[k]["dts"]:
                    rc, _, _, _, _ = m.simulate(om.default_params(max_dt=h, min_dt=h, newton_tol=1e-7), 1, h,
                                                keep_history=False)
                    if rc != 0:
                        raise SystemExit("replay of request %d failed rc=%d" % (k, rc))
                pv = m.get_solution()[0]
                m.close()
            pp, ps = pv[1::3], res[k]["pv"][1::3]
            errs.append(float(np.max(np.abs(ps - pp) / np.abs(pp))))
        out["plain_fp64"] = {"requests": n_plain, "wall_s": round(pw, 3), "solves_per_s": round(n_plain / pw, 4)}
        out["speedup_vs_plain_fp64"] = round(out["solves_per_s"] / (n_plain / pw), 3)
        out["max_rel_pressure_diff_vs_plain_fp64"] = max(errs)
        out["plain_replayed_with_served_steps"] = replayed
    return out


def copy_calibration(mib):
    """One device-to-device copy of a known size (the runtime's blit kernel):
    the reference point that scales the FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes as C
    try:
        hip = C.CDLL("libamdhip64.so")
    except OSError:
        hip = C.CDLL("/opt/rocm/lib/libamdhip64.so")
    n = int(mib) << 20
    a, b = C.c_void_p(), C.c_void_p()
    for buf in (a, b):
        if hip.hipMalloc(C.byref(buf), C.c_size_t(n)) != 0:
            raise SystemExit("hipMalloc failed")
    hip.hipMemset(a, 1, C.c_size_t(n))
    hip.hipDeviceSynchronize()
    hip.hipMemcpy(b, a, C.c_size_t(n), 3)   # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)


def beside_runs(om, np, config, dt, steps=2):
    """Newton iterations/s of workloads reported beside the headline: the
    432k-cell black-oil grid (120x120x30) in FP64, and the headline grid with
    the FP32 inner solve inside FP64 refinement, with the natural-order and
    with the two-colour tile ILU(0), in FP64 with the two-colour ILU(0), and
    both grids in FP64 with the aggregation AMG (the ParallelAmgBackend form).
    One untimed step each first, from the same initial state as the headline;
    the error of each against the FP64 natural-order run (the headline's
    arithmetic) is the largest relative pressure difference after the timed
    steps.  ILU-apply kernel time from HIP events on every 4th launch."""
    out = {}
    runs = [("c3_fp64", "c3", 0, 0, 0), ("%s_fp32_inner" % config, config, 1, 0, 0),
            ("%s_fp32_inner_colored" % config, config, 1, 1, 0), ("%s_fp64_colored" % config, config, 0, 1, 0),
            ("c3_fp32_inner_colored", "c3", 1, 1, 0), ("%s_fp64_amg" % config, config, 0, 0, 2),
            ("c3_fp64_amg", "c3", 0, 0, 2), ("%s_fp32_inner_amg" % config, config, 1, 0, 2),
            ("c3_fp32_inner_amg", "c3", 1, 0, 2)]
    ref = {}
    for key, cfg_name, precision, order, precond in [("c3_fp64", "c3", 0, 0, 0)] + runs + [("ref", config, 0, 0, 0)]:
        if key in out:
            continue
        note("beside: " + key)
        cfg = CONFIGS[cfg_name]
        case = om.BlackOilCase(om.make_spec(cfg["nx"], cfg["ny"], cfg["nz"], seed=2024, **cfg.get("spec", {})))
        prm = om.default_params(max_dt=dt, min_dt=dt, precision=precision, ilu_order=order, autotune=1,
                                precond=precond)
        model = om.Model(case)
        model.set_solution(case.pv0, case.mean0)
        rc, _, _, _, _ = model.simulate(prm, 1, dt, keep_history=False)
        model.set_profiling(4)
        t0 = time.time()
        rc2, _, _, _, st = model.simulate(prm, steps, dt, keep_history=False)
        wall = time.time() - t0
        pv = model.get_solution()[0]
        kt = model.kernel_times()
        ab = algorithmic_bytes(model.N, (model.nnzb - model.N) // 2, model.nnzb, s=8 if precision == 0 else 4,
                               colored=order == 1)
        if precond == 2:
            ab["ilu_apply"] = amg_bytes(amg_levels(cfg["nx"], cfg["ny"], cfg["nz"]), s=8 if precision == 0 else 4)[0]
        model.close()
        if key == "ref" or (precision == 0 and order == 0 and precond == 0):
            ref[cfg_name] = pv
        if key == "ref":
            continue
        ms, cnt = kt["ilu_apply"]
        out[key] = {"workload": cfg["label"], "precision": "fp64" if precision == 0 else "fp32-inner/fp64-outer",
                    "preconditioner": ("aggregation AMG" if precond == 2 else
                                       "block-ILU0, " + ("natural order" if order == 0 else "two-colour 4x2x8 tiles")),
                    "newton_iters_per_s": round(st.newton_iters / wall, 3),
                    "newton_iters": st.newton_iters,
                    "linear_iters_per_newton": round(st.linear_iters / max(st.newton_iters, 1), 2),
                    "precond_apply_avg_us": round(ms / max(cnt, 1) * 1e3, 1),
                    "precond_apply_achieved_GBs": round(ab["ilu_apply"] / max(ms / max(cnt, 1) * 1e-3, 1e-12) / 1e9,
                                                        1),
                    "steps": steps, "ok": rc == 0 and rc2 == 0, "_pv": pv, "_cfg": cfg_name}
    for key, v in out.items():
        pv = v.pop("_pv")
        r = ref.get(v.pop("_cfg"))
        if r is not None:
            v["max_rel_pressure_diff_vs_fp64_natural"] = float(np.max(np.abs(pv[1::3] - r[1::3]) / np.abs(r[1::3])))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="500k", choices=sorted(CONFIGS))
    ap.add_argument("--precision", type=int, default=0)
    ap.add_argument("--precond", type=int, default=0, choices=[0, 2],
                    help="0: block ILU(0) (the reference's default backend); 2: the aggregation AMG (FP64)")
    ap.add_argument("--ilu-order", type=int, default=0, choices=[0, 1],
                    help="block ILU(0) order of the timed run: 0 natural (the reference's), 1 two-colour tiles")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the 432k workload and the mixed-precision run reported beside the headline")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run one device copy of MIB MiB (reference for HBM byte counters)")
    ap.add_argument("--serve-requests", type=int, default=32)
    ap.add_argument("--serve-rate", type=float, default=4.0)
    ap.add_argument("--serve-steps", type=int, default=10)
    ap.add_argument("--serve-plain", type=int, default=32)
    ap.add_argument("--serve-progress", action="store_true", help="a line per finished request (long streams)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the reservoir would be split "
              "by domain decomposition; see DESIGN.md). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import omodels_amd as om

    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)

    cfg = CONFIGS[args.config]
    spec = om.make_spec(cfg["nx"], cfg["ny"], cfg["nz"], seed=2024, **cfg.get("spec", {}))
    case = om.BlackOilCase(spec)
    dt = 86400.0
    prm = om.default_params(max_dt=dt, min_dt=0.0 if cfg.get("adaptive") else dt, precision=args.precision,
                            autotune=1, ilu_order=args.ilu_order, precond=args.precond)

    t_setup = time.time()
    model = om.Model(case)
    model.set_solution(case.pv0, case.mean0)
    t_setup = time.time() - t_setup
    N, nnzb = model.N, model.nnzb
    F = (nnzb - N) // 2

    # warmup (untimed), then exactly K time steps
    if args.warmup > 0:
        rc, _, _, _, _ = model.simulate(prm, args.warmup, dt, keep_history=False)
        if rc != 0:
            raise SystemExit("warmup failed rc=%d" % rc)
    # HIP events around every 4th launch of each kernel class (an event record
    # idles the stream for a few microseconds)
    model.set_profiling(4)
    t0 = time.time()
    rc, _, dts, log, st = model.simulate(prm, args.steps, dt, keep_history=False)
    wall = time.time() - t0
    if rc != 0:
        raise SystemExit("timed run failed rc=%d" % rc)
    kt = model.kernel_times()
    # the operator application runs inside the backward sweep of the tiled ILU
    # (no launches of its own): the class then carries both kernels' bytes
    fused = kt["spmv"][1] == 0 and kt["ilu_apply"][1] > 0
    if fused:
        kt = {("ilu_apply_spmv" if k == "ilu_apply" else k): v for k, v in kt.items() if k != "spmv"}

    newton = st.newton_iters
    value = newton / wall
    ab = algorithmic_bytes(N, F, nnzb, s=8 if args.precision == 0 else 4, colored=args.ilu_order == 1)
    if args.precond == 2:
        # the preconditioner classes hold the V-cycle and the hierarchy's setup
        ab["ilu_apply"], ab["ilu_factor"] = amg_bytes(amg_levels(cfg["nx"], cfg["ny"], cfg["nz"]),
                                                      s=8 if args.precision == 0 else 4)
    kernels = {}
    for k, (ms, cnt) in kt.items():
        if cnt == 0 or k not in ab:
            continue
        avg = ms / cnt
        gbs = ab[k] / (avg * 1e-3) / 1e9
        name = {"ilu_apply": "amg_vcycle", "ilu_factor": "amg_setup"}.get(k, k) if args.precond == 2 else k
        kernels[name] = dict(total_ms=round(ms, 3), launches=cnt, avg_us=round(avg * 1e3, 2),
                             bytes_per_launch=ab[k], achieved_GBs=round(gbs, 1))
    dom = max(kernels, key=lambda k: kernels[k]["total_ms"])
    asm = kernels.get("assemble")
    roof = {"bound": "hbm", "kernel": dom, "achieved": kernels[dom]["achieved_GBs"], "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(kernels[dom]["achieved_GBs"] / HBM_PEAK_GBS, 4),
            "traffic": None}
    if args.config == "500k" and args.precision == 0 and args.ilu_order == 0 and args.precond == 0:
        roof["traffic"] = HBM_TRAFFIC_500K.get(dom)
    elif args.config == "500k" and args.precision == 1 and args.ilu_order == 1:
        roof["traffic"] = HBM_TRAFFIC_500K_FP32_COLORED.get(dom)

    line = {
        "metric": "Newton steps/s & assembly Mcells/s at 500k blackoil; ensemble p50/p99 latency",
        "value": round(value, 3),
        "unit": "newton_iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == 0 else "f64/f32",
        "data": "synthetic",
        "config": {"workload": cfg["label"], "cells": N, "nnzb": nnzb,
                   "dt_days": "adaptive, at most 1.0" if cfg.get("adaptive") else 1.0,
                   "precond": ("aggregation AMG (3x3x3 boxes, V-cycle, 2+2 Gauss-Seidel sweeps)"
                               if args.precond == 2 else
                               "block-ILU0" + (", two-colour 4x2x8 tiles" if args.ilu_order else "")),
                   "precision": "fp64" if args.precision == 0 else "fp32-inner/fp64-outer",
                   "parallelism": "single GPU"},
        "assembly_mcells_per_s": round(N * asm["launches"] / (asm["total_ms"] * 1e-3) / 1e6, 1) if asm else None,
        "newton_iters": newton,
        "linear_iters_per_newton": round(st.linear_iters / max(newton, 1), 2),
        "phase_s": {"linearize": round(st.t_linearize, 4), "solve": round(st.t_solve, 4),
                    "update": round(st.t_update, 4), "total": round(st.t_total, 4)},
        "kernels": kernels,
        "roofline": roof,
        "setup_s": round(t_setup, 3),
    }

    # FLOWS / FLORES of the final state (TpfaLinearizer::updateFlowsInfo), after
    # the timed region: kernel time from the same HIP events (class "other")
    ms0, n0 = model.kernel_times()["other"]
    for _ in range(8):
        model.flows()
    ms1, n1 = model.kernel_times()["other"]
    if n1 > n0:
        favg = (ms1 - ms0) / (n1 - n0)
        # per cell: primaries, meaning, three transmissibilities, row pointer read;
        # FLOWS and FLORES written for each neighbour entry (3 doubles each)
        fb = N * (24 + 1 + 24 + 4) + 2 * (nnzb - N) * 24
        line["flows"] = {"avg_us": round(favg * 1e3, 2), "bytes_per_launch": fb,
                         "achieved_GBs": round(fb / (favg * 1e-3) / 1e9, 1)}
    # the output quantities of the VTK files (18 fields per cell)
    for _ in range(8):
        model.output_fields()
    ms2, n2 = model.kernel_times()["other"]
    if n2 > n1:
        oavg = (ms2 - ms1) / (n2 - n1)
        ob = N * (24 + 1 + 8) + 18 * N * 8
        line["output_fields"] = {"avg_us": round(oavg * 1e3, 2), "bytes_per_launch": ob,
                                 "achieved_GBs": round(ob / (oavg * 1e-3) / 1e9, 1)}
    # the assembly with molecular diffusion in the oil phase (the diffusive face
    # term), at the final state: same HIP events (class "assemble")
    ms3, n3 = model.kernel_times()["assemble"]
    model.set_diffusion(3e-4, 2e-4)
    for _ in range(8):
        om.check(om.lib().omodels_linearizer_linearize_domain(model.h, dt, 1), "linearize_domain")
    ms4, n4 = model.kernel_times()["assemble"]
    if n4 > n3:
        davg = (ms4 - ms3) / (n4 - n3)
        db = ab["assemble"] + F * 8          # plus one diffusivity per face
        line["assemble_diffusion"] = {"avg_us": round(davg * 1e3, 2), "bytes_per_launch": db,
                                      "achieved_GBs": round(db / (davg * 1e-3) / 1e9, 1),
                                      "mcells_per_s": round(N / (davg * 1e-3) / 1e6, 1)}
    # and with dispersion as well (dispersion coefficient 10 m in every cell)
    model.set_dispersion(cell_dispersion=np.full(N, 10.0))
    for _ in range(8):
        om.check(om.lib().omodels_linearizer_linearize_domain(model.h, dt, 1), "linearize_domain")
    ms5, n5 = model.kernel_times()["assemble"]
    model.set_dispersion(off=True)
    model.set_diffusion(None)
    if n5 > n4:
        eavg = (ms5 - ms4) / (n5 - n4)
        # plus a dispersivity per face, the cells' velocities read and written
        eb = ab["assemble"] + 2 * F * 8 + 2 * N * 8
        line["assemble_diffusion_dispersion"] = {"avg_us": round(eavg * 1e3, 2), "bytes_per_launch": eb,
                                                 "achieved_GBs": round(eb / (eavg * 1e-3) / 1e9, 1),
                                                 "mcells_per_s": round(N / (eavg * 1e-3) / 1e6, 1)}
    if not args.no_extra:
        note("headline done: %.3f newton_iters/s" % value)
        line["beside"] = beside_runs(om, np, args.config, dt)
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc
        prob = orc.problem_from_case(case)
        threads = min(16, os.cpu_count() or 1)
        oprm = orc.params(max_dt=dt, min_dt=dt)
        t0 = time.time()
        rco, pvo, mo, histo, dtso, logo, tim = orc.simulate(prob, case.pv0, case.mean0, args.cpu_steps, dt, oprm,
                                                            nthreads=threads, keep_history=False)
        cw = time.time() - t0
        # every record but the converged check that ends each time step is a Newton update
        cn = len(logo) - args.cpu_steps
        # the assembly alone, median of 5 linearizations on all threads used
        # and on one (the linear solve is sequential in both)
        st1 = orc.storage(prob, case.pv0, case.mean0)

        def asm_rate(nt):
            ts = orc.linearize_times(prob, case.pv0, case.mean0, st1, dt, nt, reps=5)
            return round(N / float(np.median(ts)) / 1e6, 2)

        cpu_model = ""
        try:
            with open("/proc/cpuinfo") as f:
                for ln in f:
                    if ln.startswith("model name"):
                        cpu_model = ln.split(":", 1)[1].strip()
                        break
        except OSError:
            pass
        line["cpu_baseline"] = {
            "value": round(cn / cw, 4), "unit": "newton_iters/s", "cores": threads, "kind": "port",
            "sample": "%d time step(s) of the same 500k workload from the initial state, oracle (C++ restatement): "
                      "OpenMP assembly on %d threads, sequential ILU0/BiCGStab as the reference" % (args.cpu_steps, threads),
            "wall_s": round(cw, 2), "newton_iters": cn,
            "assembly_mcells_per_s": asm_rate(threads),
            "assembly_mcells_per_s_1_thread": asm_rate(1),
            "nproc": os.cpu_count(), "cpu_model": cpu_model,
        }
    if not args.no_serve:
        note("cpu baseline done")
        line["ensemble"] = serve_workload(om, np, args.serve_requests, args.serve_rate, args.serve_steps,
                                          min(args.serve_plain, args.serve_requests),
                                          progress=args.serve_progress)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
