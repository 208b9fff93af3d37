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

  python bench.py [--steps K] [--warmup W] [--gpus 1] [--config 500k|c2|c3]
                  [--precision 0|1] [--no-cpu]
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
}


def algorithmic_bytes(N, F, nnzb, n=3, s=8):
    """Per-launch algorithmic HBM bytes of each kernel class."""
    return {
        # primaries, old storage, porosity, depth, meaning flags; one T per face;
        # residual and Jacobian writes
        "assemble": N * (2 * n * 8 + 8 + 8 + 4) + F * 8 + N * n * 8 + nnzb * n * n * s,
        "spmv": nnzb * (n * n * s + 4) + (N + 1) * 4 + 2 * N * n * s,
        "ilu_apply": nnzb * (n * n * s + 4) + (N + 1) * 4 + 3 * N * n * s,
        "ilu_factor": 2 * nnzb * n * n * 8 + nnzb * 4 + (N + 1) * 4,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="500k", choices=sorted(CONFIGS))
    ap.add_argument("--precision", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=1)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the reservoir would be split "
              "by domain decomposition; see DESIGN.md). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import omodels_amd as om

    cfg = CONFIGS[args.config]
    spec = om.make_spec(cfg["nx"], cfg["ny"], cfg["nz"], seed=2024)
    case = om.BlackOilCase(spec)
    dt = 86400.0
    prm = om.default_params(max_dt=dt, min_dt=dt, precision=args.precision)

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
    model.set_profiling(True)
    t0 = time.time()
    rc, _, dts, log, st = model.simulate(prm, args.steps, dt, keep_history=False)
    wall = time.time() - t0
    if rc != 0:
        raise SystemExit("timed run failed rc=%d" % rc)
    kt = model.kernel_times()

    newton = st.newton_iters
    value = newton / wall
    ab = algorithmic_bytes(N, F, nnzb)
    kernels = {}
    for k, (ms, cnt) in kt.items():
        if cnt == 0 or k not in ab:
            continue
        avg = ms / cnt
        gbs = ab[k] / (avg * 1e-3) / 1e9
        kernels[k] = dict(total_ms=round(ms, 3), launches=cnt, avg_us=round(avg * 1e3, 2),
                          bytes_per_launch=ab[k], achieved_GBs=round(gbs, 1))
    dom = max(kernels, key=lambda k: kernels[k]["total_ms"])
    asm = kernels.get("assemble")
    roof = {"bound": "hbm", "kernel": dom, "achieved": kernels[dom]["achieved_GBs"], "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(kernels[dom]["achieved_GBs"] / HBM_PEAK_GBS, 4), "traffic": None}

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
        "config": {"workload": cfg["label"], "cells": N, "nnzb": nnzb, "dt_days": 1.0,
                   "precond": "block-ILU0", "precision": "fp64" if args.precision == 0 else "fp32-inner/fp64-outer",
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
        line["cpu_baseline"] = {
            "value": round(cn / cw, 4), "unit": "newton_iters/s", "cores": threads, "kind": "port",
            "sample": "%d time step(s) of the same 500k workload from the initial state, oracle (C++ restatement): "
                      "OpenMP assembly on %d threads, sequential ILU0/BiCGStab as the reference" % (args.cpu_steps, threads),
            "wall_s": round(cw, 2), "newton_iters": cn,
            "assembly_mcells_per_s": round(N * (cn + args.cpu_steps) / max(tim[0], 1e-9) / 1e6, 2),
        }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
