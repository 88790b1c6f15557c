#!/usr/bin/env python3
"""Benchmark of the quasi-dynamic boundary-integral hot path on one MI355X.

Metric: simulated years per second per GPU, batched vs plain binary64, at
the binary64 slip error.

Workload (default, "ensemble"): 256 independent fault instances in 4
geometry groups of 64 (1024, 2048, 4096, 8192 Okada rectangles of 2 km, the
30-degree thrust layout), per-instance friction draws and seeded initial
states.  Each group shares one K; every RK45 stage of every instance is one
batched matvec per group inside a captured hipGraph.  A "step" is one pass of
all 256 instances through the simulated window from the same initial states.

Reported:
  value         batched binary64 (K on v_mfma_f64_16x16x4f64): aggregate
                simulated years per wall second; its per-instance arithmetic
                and step decisions are those of the plain path (binary64 test
                parity), so it meets the binary64 slip-error bar.
  mixed         binary64 step-start matvec t_0 = K v_0 and split-bf16 (hi+lo,
                fp32 MFMA accumulation) stage increments K (v_i - v_0) on the
                same window: aggregate sim-yr/s and slip error against the
                batched binary64 result.
  split_bf16    as mixed with t_0 in split bf16 too.
  plain_fp64    the single-instance device path (streaming binary64 GEMV per
                instance): measured on instance 0 of every group, costed per
                attempt over all instances of the group.
  roofline      the stage GEMM of the batched binary64 run (dominant kernel),
                timed with HIP events on the library stream.
  cpu_baseline  the oracle's C restatement (OpenMP over the host cores): RK45
                attempts/s on instance 0 of every group over a bounded sample,
                costed over all instances with the GPU run's attempt counts.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--years Y]
Prints exactly one JSON line on stdout; progress goes to stderr.
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
FP64_MFMA_PEAK = 78.6e12   # FLOP/s, dense binary64 MFMA
BF16_MFMA_PEAK = 2.5e15    # FLOP/s, dense bf16 MFMA


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cores() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except Exception:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def _progress(tag):
    def f(st):
        log(f"[{tag}] attempts {st.attempts} steps {st.steps} t_min {st.min_time:.4g} s "
            f"t_max {st.max_time:.4g} s failed {st.n_failed} wall {st.wall_seconds:.1f} s")
    return f


def run_window(ens, T, tag, chunk=2000):
    ens.reset()
    st = ens.run(T, max_attempts=10**7, chunk_attempts=chunk, progress=_progress(tag))
    return st


def gemm_shapes(info):
    return [(3 * m.n, 2 * m.n, y0.shape[0]) for m, d, y0 in info]


def bench_ensemble(args):
    from unicycle_amd import Engine, PREC_F64, PREC_MIXED, PREC_SPLIT_BF16, configs
    T = args.years * configs.YEAR
    eng = Engine(0)
    t0 = time.perf_counter()
    ens, info = configs.ensemble_setup(eng, PREC_F64)
    t_setup = time.perf_counter() - t0
    shapes = gemm_shapes(info)
    n_inst = sum(s[2] for s in shapes)
    log(f"ensemble setup+build {t_setup:.1f} s; groups {[(m, n) for m, n, _ in shapes]}")
    for w in range(args.warmup):
        run_window(ens, min(T, 1e-3 * configs.YEAR), f"warmup {w}")
    walls, st = [], None
    t_start = time.perf_counter()
    for k in range(args.steps):
        st = run_window(ens, T, f"f64 step {k}")
        walls.append(st.wall_seconds)
    t_total = time.perf_counter() - t_start
    if st.n_failed:
        raise RuntimeError(f"{st.n_failed} instances failed in the window")
    value = st.inst_time_sum / configs.YEAR / (t_total / args.steps)
    y64 = [ens.state(g) for g in range(len(info))]
    # dominant kernel: the stage GEMM over all groups, HIP events on its stream
    t_gemm, gemm_bytes = ens.time_gemm(20)
    flops = sum(2.0 * m * n * 64 for m, n, _ in shapes)
    achieved = flops / t_gemm
    kb, _ = ens.tuning()
    ens.close()

    # reduced-precision stage matvecs on the same window
    reduced = {}
    for name, prec in (("mixed", PREC_MIXED), ("split_bf16", PREC_SPLIT_BF16)):
        ensr, _ = configs.ensemble_setup(eng, prec)
        if args.warmup:
            run_window(ensr, min(T, 1e-3 * configs.YEAR), f"{name} warmup")
        str_ = run_window(ensr, T, name)
        yr = [ensr.state(g) for g in range(len(info))]
        t_g, b_g = ensr.time_gemm(20)
        ensr.close()
        num = den = 0.0
        for (ya, *_), (yb, *_) in zip(y64, yr):
            sa = np.hypot(ya[:, 0::7], ya[:, 1::7])
            sb = np.hypot(yb[:, 0::7], yb[:, 1::7])
            num += float(np.sum((sb - sa) ** 2))
            den += float(np.sum(sa ** 2))
        reduced[name] = {
            "value": str_.inst_time_sum / configs.YEAR / str_.wall_seconds, "unit": "sim-yr/s",
            "slip_error_vs_f64": (num / den) ** 0.5, "failed": int(str_.n_failed),
            "attempts": int(str_.attempts), "wall_s": str_.wall_seconds,
            "stage_gemm_us": 1e6 * t_g, "stage_gemm_GBps": b_g / t_g / 1e9,
            "stage_gemm_hbm_frac": b_g / t_g / HBM_PEAK,
            "stage_gemm_useful_TFLOPs": flops / t_g / 1e12}
    reduced["mixed"]["dtype"] = "binary64 K for t_0 = K v_0, split-bf16 increments, f64 state"
    reduced["split_bf16"]["dtype"] = "split-bf16 K (t_0 and increments), f64 state"

    # plain binary64 device path on instance 0 of each group
    plain_time = 0.0
    plain_detail = []
    cpu_rates = []
    for gi, (m, d, y0) in enumerate(info):
        p = Engine(0)
        p.set_elastic(m.lambda_, m.mu)
        p.set_rectangles(*m.geometry())
        p.set_friction(*configs.instance_friction(m, d, 0))
        p.buildG()
        res = p.integrate(y0[0], T)
        att = y64[gi][4]
        per_att = res.wall_seconds / res.attempts
        plain_time += per_att * float(np.sum(att))
        plain_detail.append({"patches": m.n, "attempts_inst0": int(res.attempts),
                             "us_per_attempt": 1e6 * per_att,
                             "ensemble_attempts_inst0": int(att[0])})
        log(f"plain group {gi}: {res.attempts} attempts, {1e6 * per_att:.0f} us/attempt")
        if not args.no_cpu:
            cpu_rates.append((gi, p.get_kernel(), m, d, y0[0]))
        p.close()
    value_plain = n_inst * T / configs.YEAR / plain_time

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "sim-yr/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_total / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded friction draws and initial slip-rate perturbations)",
        "config": {"workload": "batched ensemble: 4 groups x 64 instances of 1024/2048/4096/8192 "
                               f"2-km Okada patches (30 deg thrust), window {args.years:g} yr per "
                               "instance, binary64 K on MFMA, RK45 eps 1e-6",
                   "instances": n_inst, "groups": [[m, n] for m, n, _ in shapes],
                   "window_s": T, "gemm_kb_per_split": kb},
        "vs_plain_fp64": value / value_plain,
        "plain_fp64": {"value": value_plain, "unit": "sim-yr/s",
                       "method": "single-instance device path per instance, measured on "
                                 "instance 0 of each group, costed per attempt",
                       "groups": plain_detail},
        "mixed": reduced["mixed"],
        "split_bf16": reduced["split_bf16"],
        "roofline": {"bound": "mfma", "achieved": achieved / 1e12, "peak": FP64_MFMA_PEAK / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / FP64_MFMA_PEAK, "traffic": None,
                     "kernel": "ucy::gemm_kernel (binary64 MFMA stage GEMM, all groups)",
                     "avg_launch_us": 1e6 * t_gemm, "flops_per_launch": flops,
                     "bytes_per_launch": gemm_bytes,
                     "hbm_GBps": gemm_bytes / t_gemm / 1e9},
        "detail": {"attempts": int(st.attempts), "accepted_steps": int(st.steps),
                   "replays": int(st.replays), "wall_s": walls,
                   "min_time_s": st.min_time, "max_time_s": st.max_time,
                   "rhs_per_s": 6.0 * st.attempts / (t_total / args.steps),
                   "setup_s": t_setup},
    }
    return out, (cpu_rates, y64, T)


def cpu_baseline_ensemble(cpu_inputs, y64, T, budget_s=20.0):
    """Oracle (C restatement, OpenMP) on the host cores: attempts/s of
    instance 0 of each group over a bounded sample of its run, costed over
    every instance's attempt count of the GPU run."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import orc
    from unicycle_amd import configs
    cores = host_cores()
    total, samples = 0.0, []
    n_inst = 0
    for gi, K, m, d, y0 in cpu_inputs:
        sysm = orc.System(configs.instance_friction(m, d, 0), K, nthreads=cores)
        t0 = time.perf_counter()
        sysm.odefun(y0)
        t_rhs = time.perf_counter() - t0
        n_steps = max(2, int(budget_s / len(cpu_inputs) / (6.5 * t_rhs)))
        t0 = time.perf_counter()
        _, st, *_ = sysm.integrate(y0, 1e30, max_steps=n_steps, log_cap=0)
        wall = time.perf_counter() - t0
        rate = st.attempts / wall
        att = y64[gi][4]
        n_inst += att.size
        total += float(np.sum(att)) / rate
        samples.append(f"{m.n} patches: {st.attempts} attempts in {wall:.1f} s")
        log(f"cpu group {gi}: {rate:.2f} attempts/s")
    return {"value": n_inst * T / configs.YEAR / total, "unit": "sim-yr/s", "cores": cores,
            "kind": "port",
            "sample": "instance 0 of each group from its initial state, " + "; ".join(samples)
                      + "; costed with every instance's GPU attempt count over the window"}


def bench_plain(args):
    """Single instance, binary64: the 2048-patch thrust (64 x 32 rectangles)."""
    from unicycle_amd import Engine, configs
    model = configs.thrust_fault(nx=64, nz=32)
    eng = Engine(0)
    eng.set_elastic(model.lambda_, model.mu)
    eng.set_rectangles(*model.geometry())
    eng.set_friction(*model.friction())
    eng.buildG()
    m, n = eng.shape
    y0 = configs.perturbed_initial_state(model, eng.initStateVector(), seed=1)
    T = args.years * configs.YEAR
    for _ in range(args.warmup):
        eng.integrate(y0, min(T, 0.01 * configs.YEAR))
    t_start = time.perf_counter()
    for _ in range(args.steps):
        res = eng.integrate(y0, T)
    t_total = time.perf_counter() - t_start
    t_k = eng.time_rhs_kernel(100)
    algo_bytes = 8.0 * m * n + 8.0 * (n + m) + 7 * 8.0 * (m // 3) * 2
    return {
        "metric": METRIC, "value": res.time / configs.YEAR / (t_total / args.steps),
        "unit": "sim-yr/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * t_total / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded thrust fault, perturbed steady state)",
        "config": {"workload": f"plain binary64, 2048-patch thrust, window {args.years:g} yr"},
        "roofline": {"bound": "hbm", "achieved": algo_bytes / t_k / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": algo_bytes / t_k / HBM_PEAK, "traffic": None,
                     "kernel": "ucy::gemv_post_kernel", "avg_launch_us": 1e6 * t_k},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--years", type=float, default=0.01)
    ap.add_argument("--workload", default="ensemble", choices=["ensemble", "plain"])
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build runs on a single MI355X and has no "
              f"multi-GPU path (--gpus {args.gpus} requested)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "plain":
        print(json.dumps(bench_plain(args)), flush=True)
        return
    out, (cpu_inputs, y64, T) = bench_ensemble(args)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_ensemble(cpu_inputs, y64, T)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
