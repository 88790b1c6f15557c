#!/usr/bin/env python3
"""Benchmark of the quasi-dynamic boundary-integral hot path on one MI355X.

Metric: simulated years per second per GPU (batched-mixed vs plain FP64).

A "step" is one pass of the hot path over one batch of synthetic input: one
integration of the benchmark window (the same seeded initial state every
step, so every step does identical work).

Workloads
  plain     the 2048-patch thrust (64 x 32 Okada rectangles of 2 km, 30 deg dip,
            a=0.010, b=0.014 in the 40x20 km ellipse / 0.006 outside), FP64 K,
            device-resident Cash-Karp RK45; value = simulated yr/s.

The CPU baseline times the oracle's C restatement of the same algorithm
(OpenMP over the host cores) on a bounded sample and extrapolates with the
GPU run's attempts per simulated year (identical accept/reject decisions).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--years Y]
Prints exactly one JSON line on stdout.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "simulated yr/s per GPU, batched-mixed vs plain FP64, at ≤1e-6 slip error"
HBM_PEAK = 8.0e12          # B/s, MI355X HBM3E
FP64_PEAK = 78.6e12        # FLOP/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cores() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except Exception:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def cpu_baseline_plain(model, K, y0, attempts_per_year, budget_s=15.0):
    """Oracle (C restatement, OpenMP) on the host cores: attempts/s over a
    bounded sample of the same integration, converted to sim-yr/s with the GPU
    run's attempts per simulated year."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import orc
    cores = host_cores()
    sysm = orc.System(model.friction(), K, nthreads=cores)
    # one timed RHS first to size the sample
    t0 = time.perf_counter()
    sysm.odefun(y0)
    t_rhs = time.perf_counter() - t0
    n_steps = max(3, int(budget_s / (6.5 * t_rhs)))
    t0 = time.perf_counter()
    _, st, tl, dl, rc = sysm.integrate(y0, 1e30, max_steps=n_steps, log_cap=0)
    wall = time.perf_counter() - t0
    att_per_s = st.attempts / wall
    return {
        "value": att_per_s / attempts_per_year,
        "unit": "sim-yr/s",
        "cores": cores,
        "kind": "port",
        "sample": (f"{st.attempts} RK45 attempts ({st.rhs_evals} RHS) of the same run from the "
                   f"same initial state in {wall:.1f} s ({att_per_s:.2f} attempts/s), "
                   f"converted with the GPU window's {attempts_per_year:.0f} attempts/yr"),
        "attempts_per_s": att_per_s,
    }


def bench_plain(args):
    from unicycle_amd import Engine, configs
    model = configs.thrust_fault(nx=64, nz=32)
    eng = Engine(0)
    eng.set_elastic(model.lambda_, model.mu)
    eng.set_rectangles(*model.geometry())
    eng.set_friction(*model.friction())
    t0 = time.perf_counter()
    eng.buildG()
    t_build = time.perf_counter() - t0
    m, n = eng.shape
    y0 = configs.perturbed_initial_state(model, eng.initStateVector(), seed=1)
    T = args.years * configs.YEAR
    for _ in range(args.warmup):
        eng.integrate(y0, min(T, 0.2 * configs.YEAR))
    walls, res = [], None
    t_start = time.perf_counter()
    for _ in range(args.steps):
        res = eng.integrate(y0, T)
        walls.append(res.wall_seconds)
    t_total = time.perf_counter() - t_start
    ms_per_step = 1e3 * t_total / args.steps
    sim_years = res.time / configs.YEAR
    value = sim_years / (t_total / args.steps)
    # dominant kernel: fused matvec + post-pass, timed live with HIP events
    t_k = eng.time_rhs_kernel(100)
    algo_bytes = 8.0 * m * n + 8.0 * (n + m) + 7 * 8.0 * (m // 3) * 2
    achieved = algo_bytes / t_k
    attempts_per_year = res.attempts / sim_years
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "sim-yr/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded thrust fault, perturbed steady state)",
        "config": {"workload": "plain FP64, 2048-patch thrust (64x32 Okada, 2 km, 30 deg), "
                               f"window {args.years:g} yr from t=0",
                   "patches": model.n, "K_shape": [m, n], "epsilon": 1e-6},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": None,
                     "kernel": "ucy::gemv_post_kernel (K v + odefun post-pass)",
                     "avg_launch_us": 1e6 * t_k, "bytes_per_launch": algo_bytes},
        "detail": {"rhs_per_s": res.rhs_evals / res.wall_seconds,
                   "attempts": res.attempts, "accepted_steps": res.steps,
                   "rhs_evals": res.rhs_evals, "us_per_attempt": 1e6 * res.wall_seconds / res.attempts,
                   "buildG_s": t_build, "assembly_pairs_per_s": model.n ** 2 / t_build},
    }
    return out, (model, eng, y0, attempts_per_year)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--years", type=float, default=4.0)
    ap.add_argument("--workload", default="plain", choices=["plain"])
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build runs on a single MI355X and has no "
              f"multi-GPU path (--gpus {args.gpus} requested)", file=sys.stderr)
        sys.exit(2)
    out, extra = bench_plain(args)
    if not args.no_cpu:
        model, eng, y0, apy = extra
        K = eng.get_kernel()
        out["cpu_baseline"] = cpu_baseline_plain(model, K, y0, apy)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
