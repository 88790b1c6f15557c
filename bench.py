#!/usr/bin/env python3
"""Benchmark of the quasi-dynamic boundary-integral hot path on one MI355X.

Metric: simulated years per second per GPU, batched vs plain binary64.

Workload (default, "ensemble"): 256 independent fault instances in 4
geometry groups of 64 (1024, 2048, 4096, 8192 Okada rectangles of 2 km, the
30-degree thrust layout), per-instance friction draws and seeded initial
states.  Each group shares one K; every RK45 stage of every instance is one
batched matvec per group inside a captured hipGraph.

A "step" is --attempts RK45 attempts of every instance.  The headline
variant runs --warmup untimed steps and then --steps timed steps, all
continuing ONE run from the seeded initial states, so the timed window
starts after the start-up growth of the step size; ms_per_step is the mean
wall time of the timed steps (each bracketed by device synchronisation) and
value is the simulated time the 256 instances advance in them, summed, per
wall second.

Sections, in order (the first five form the line; the rest run only while
the --budget wall clock allows, and a failing section becomes an "errors"
entry instead of a lost line):
  batched_f64   binary64 K on v_mfma_f64_16x16x4f64, --attempts attempts
                from the initial states, timed once; it is the precision
                reference: slip at a fixed pre-event time (every instance
                clamped to land on it) and first-event times.
  training      the 8-64-64-1 step-size network (hand-written GPU kernels),
                trained on instances drawn apart from the bench's.
  headline      --headline (default bf16+i8 K with the learned proposals):
                the timed steps, its stage-GEMM timing (HIP events on the
                library stream) and its slip and first-event errors against
                binary64.  value is reported only if it meets the bar (slip
                <= 1e-6, events within 0.1 %, none in only one run);
                otherwise value falls back to batched binary64.
  plain_fp64    the single-instance binary64 device path (streaming GEMV),
                seconds per attempt of instance 0 of every group, costed over
                every instance's attempts of the headline window.
  cpu_baseline  the oracle's C restatement (OpenMP over the host cores) per
                attempt on instance 0 of every group over a bounded sample,
                costed the same way.
  bf16_i8, split_bf16, mixed
                the other precisions on the reference's schedule, with their
                errors against binary64.
  learned_step_proposal
                the network against the classic controller on 48 held-out
                instances over a fixed simulated window.
  assembly      Green's-function assembly (buildG) on the device for the
                three source types at the single-instance model sizes.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--attempts R]
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
INF_T = 1e30               # no end time: runs are bounded by attempts
# HBM bytes per launch of the dominant kernels from rocprofv3 --pmc FETCH_SIZE /
# WRITE_SIZE runs, scaled by a 2 GiB copy (profiles/copy_calibration_2GiB.md,
# profiles/stage_gemm_split_bf16_ensemble256.md,
# profiles/stage_gemm_bf16_i8_ensemble256.md); same workload as this bench
MEASURED_TRAFFIC = {"split_bf16": 2.269e9, "bf16_i8": 1.733e9}
# FP64 flop per (receiver, source) pair of the assembly kernels from
# rocprofv3 --pmc SQ_INSTS_VALU_FLOPS_FP64 x 64 lanes / pairs on the same
# models (profiles/assembly_greens_functions.md).  The strain-volume figure
# is the evaluation that forms every displacement-gradient entry once; the
# device kernel splits it by displacement component and executes
# EXECUTED_FLOP_PER_PAIR (ADD + MUL + 2 FMA + TRANS counters), reported beside it
ASSEMBLY_FLOP_PER_PAIR = {"okada": 7108.0, "okada_stress": 7199.0, "nikkhoo": 21680.0,
                          "strain_volume": 61800.0}
EXECUTED_FLOP_PER_PAIR = {"strain_volume": 106700.0}
FP64_VALU_PEAK = 78.6e12   # FLOP/s, dense binary64 vector


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cores() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except Exception:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def host_info() -> dict:
    """The host the CPU baseline ran on: CPU model and the cores visible."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "cores_visible": os.cpu_count(), "cores_used": host_cores()}


def _progress(tag):
    def f(st):
        log(f"[{tag}] attempts {st.attempts} steps {st.steps} t_min {st.min_time:.4g} s "
            f"t_max {st.max_time:.4g} s failed {st.n_failed} wall {st.wall_seconds:.1f} s")
    return f


class Budget:
    """Wall-clock budget of the whole bench (seconds since the process
    started): sections past the core run only while their estimate fits."""

    def __init__(self, seconds: float):
        self.t0 = time.perf_counter()
        self.limit = float(seconds)

    def elapsed(self) -> float:
        return time.perf_counter() - self.t0

    def allows(self, estimate: float) -> bool:
        return self.elapsed() + estimate <= self.limit


def section(out: dict, name: str, budget: Budget, estimate: float, fn):
    """Run fn() as an optional section of the JSON line: skipped when its
    estimate does not fit the wall budget, an exception becomes an error
    entry instead of a lost line."""
    timing = out.setdefault("section_seconds", {})
    if not budget.allows(estimate):
        out.setdefault("skipped", {})[name] = (f"wall budget: {budget.elapsed():.0f} s used of "
                                               f"{budget.limit:.0f}, estimate {estimate:.0f} s")
        log(f"section {name} skipped ({out['skipped'][name]})")
        return None
    t0 = time.perf_counter()
    try:
        return fn()
    except Exception as e:  # recorded in the line, the bench goes on
        out.setdefault("errors", {})[name] = f"{type(e).__name__}: {e}"
        log(f"section {name} failed: {type(e).__name__}: {e}")
        return None
    finally:
        timing[name] = round(time.perf_counter() - t0, 1)


def run_attempts(ens, R, tag):
    """R attempts per instance from the initial states (no end time)."""
    ens.reset()
    ens.set_stop_at_end(False)
    ens.set_events(1e-3, False)
    return ens.run(INF_T, max_attempts=R, chunk_attempts=max(1, R // 2), progress=_progress(tag))


def slip_at(ens, T, tag):
    """Every instance clamped to land exactly on T (measurement option)."""
    ens.reset()
    ens.set_stop_at_end(True)
    ens.set_events(1e-3, False)
    st = ens.run(T, max_attempts=10**6, chunk_attempts=2000, progress=_progress(tag))
    ens.set_stop_at_end(False)
    if st.n_failed or st.min_time < T:   # no slip figure: the bar then fails
        log(f"{tag}: {st.n_failed} failed, t_min {st.min_time:.4g} < {T:.4g}")
        return st, None
    return st, [ens.state(g)[0] for g in range(len(ens.groups))]


def first_events(ens, T, tag):
    """Every instance stopped at its first event (vMax > 1e-3 m/s) or at T."""
    ens.reset()
    ens.set_events(1e-3, True)
    ens.run(T, max_attempts=10**6, chunk_attempts=2000, progress=_progress(tag))
    ens.set_events(1e-3, False)
    return np.concatenate([ens.first_events(g) for g in range(len(ens.groups))])


def slip_error(ya, yb):
    if ya is None or yb is None:
        return None
    num = den = 0.0
    for a, b in zip(ya, yb):
        sa = np.hypot(a[:, 0::7], a[:, 1::7])
        sb = np.hypot(b[:, 0::7], b[:, 1::7])
        num += float(np.sum((sb - sa) ** 2))
        den += float(np.sum(sa ** 2))
    return (num / den) ** 0.5


def event_error(ta, tb):
    """First-event times tb against ta (-1: no event by the horizon).  Events
    found by one run only are counted before anything else, so a run that
    loses (or invents) every event cannot pass for one that matches."""
    a_ev, b_ev = ta > 0, tb > 0
    both = a_ev & b_ev
    out = {"instances_with_event": int(both.sum()),
           "event_only_in_one": int((a_ev ^ b_ev).sum())}
    if both.any():
        r = np.abs(tb[both] - ta[both]) / ta[both]
        out.update({"rel_err_median": float(np.median(r)), "rel_err_max": float(r.max()),
                    "within_0.1pct": float(np.mean(r <= 1e-3))})
    return out


def meets_bar(slip_err, ev) -> bool:
    """The headline bar: slip error <= 1e-6 at the pre-event time, first-event
    times within 0.1 %, no event in only one of the two runs.  A missing
    figure fails."""
    if slip_err is None or not np.isfinite(slip_err) or slip_err > 1e-6:
        return False
    if ev is None or ev.get("event_only_in_one", 1) != 0:
        return False
    if ev.get("instances_with_event", 0) > 0 and not ev.get("rel_err_max", np.inf) <= 1e-3:
        return False
    return True


TRAIN_SEED = 200   # draws of the network's training instances (the bench's are 100 + group)


def train_step_net(eng, args):
    """The step-size network for the headline ensemble, trained apart from
    the bench's instances: the plain binary64 integrator's step histories of
    16 instances of the 1024-patch geometry (draws seeded TRAIN_SEED), then
    two closed-loop rounds and the output offset on an ensemble of 16
    training instances of each of the four geometries in split bf16."""
    from unicycle_amd import Engine, PREC_SPLIT_BF16, configs, stepnet
    t0 = time.perf_counter()
    nx, nz = configs.ENSEMBLE_GROUPS[0]
    m = configs.thrust_fault(nx=nx, nz=nz, ellipse=None)
    d = configs.ensemble_instances(m, 16, seed=TRAIN_SEED)
    ic = Engine(0)
    ic.set_elastic(m.lambda_, m.mu)
    ic.set_rectangles(*m.geometry())
    y0 = []
    for i in range(16):
        ic.set_friction(*configs.instance_friction(m, d, i))
        y0.append(configs.perturbed_initial_state(m, ic.initStateVector(), int(d["seeds"][i])))
    X, T = stepnet.training_data(ic, m, d, range(16), args.learned_steps, np.array(y0))
    ic.close()
    net = stepnet.StepNet(eng, seed=0)
    net.fit_normalization(X, T)
    loss = net.train(X, T, epochs=40, batch=4096, lr=2e-3, seed=1)
    n_plain = X.shape[0]
    tr, _ = configs.ensemble_setup(eng, PREC_SPLIT_BF16, n_inst=16, seed=TRAIN_SEED)
    X, T = stepnet.fit_closed_loop(net, X, T, tr, args.learned_steps, rounds=2,
                                   progress=_progress("train closed-loop"))
    offset, tried = stepnet.calibrate_offset(net, tr, attempts=max(200, args.attempts // 2),
                                             progress=_progress("train offset"))
    tr.close()
    info = {"plain_f64_samples": n_plain, "samples": int(X.shape[0]),
            "training_instances": "16 per geometry, draws seeded apart from the bench's",
            "loss_first_last": [float(loss[0]), float(loss[-1])], "closed_loop_rounds": 2,
            "output_offset_log10": offset,
            "offset_sim_seconds": {str(k): (v if np.isfinite(v) else None) for k, v in tried.items()},
            "train_s": time.perf_counter() - t0}
    log(f"step-size network: {info}")
    return net, info


VARIANTS = {  # name: (precision constant, learned proposals, dtype text)
    "bf16_i8_learned": ("PREC_BF16_I8", True,
                        "bf16+i8 K (3 bytes per entry), split-bf16 V, fp32 MFMA accumulate, "
                        "f64 state, learned step-size proposals (fp32 MLP)"),
    "bf16_i8": ("PREC_BF16_I8", False,
                "bf16+i8 K (3 bytes per entry), split-bf16 V, fp32 MFMA accumulate, f64 state"),
    "split_bf16": ("PREC_SPLIT_BF16", False,
                   "split-bf16 K (hi+lo, 4 bytes per entry), fp32 MFMA accumulate, f64 state"),
    "mixed": ("PREC_MIXED", False,
              "binary64 K for t_0 = K v_0, split-bf16 increments, f64 state"),
}


def _prec(name):
    import unicycle_amd
    return getattr(unicycle_amd, VARIANTS[name][0])


def group_attempts(ens):
    return [ens.state(g)[4].astype(np.int64) for g in range(len(ens.groups))]


def bench_reference(args, eng, shapes_out):
    """Batched binary64 (the precision reference of every reduced variant):
    R attempts per instance from the initial states, timed once; then the
    slip at the pre-event time and the first-event times."""
    from unicycle_amd import PREC_F64, configs
    R = args.attempts
    t0 = time.perf_counter()
    ens, info = configs.ensemble_setup(eng, PREC_F64)
    setup_s = time.perf_counter() - t0
    shapes_out.extend((3 * m.n, 2 * m.n, y0.shape[0]) for m, d, y0 in info)
    log(f"ensemble setup+build {setup_s:.1f} s; groups {[(m, n) for m, n, _ in shapes_out]}")
    t0 = time.perf_counter()
    st = run_attempts(ens, R, "f64")
    wall = time.perf_counter() - t0
    if st.n_failed:
        raise RuntimeError(f"{st.n_failed} binary64 instances failed")
    t_gemm, gemm_bytes = ens.time_gemm(20)
    kb, _, _ = ens.tuning()
    _, ys = slip_at(ens, args.slip_years * configs.YEAR, "f64 slip window")
    if ys is None:
        raise RuntimeError("binary64 ensemble did not reach the slip window")
    te = first_events(ens, args.event_years * configs.YEAR, "f64 first events")
    ens.close()
    return {"info": info, "ys": ys, "te": te, "st": st, "wall": wall, "t_gemm": t_gemm,
            "gemm_bytes": gemm_bytes, "kb": kb, "setup_s": setup_s}


def bench_headline(args, eng, name, net, ref):
    """The headline variant: warmup W and then K timed steps of R attempts
    per instance each, continuing one run from the seeded initial states
    (step k covers attempts (W+k)R+1 .. (W+k+1)R of every instance), then its
    GEMM timing and its slip and first-event errors against binary64."""
    from unicycle_amd import configs
    R, W, K = args.attempts, args.warmup, args.steps
    ens, _ = configs.ensemble_setup(eng, _prec(name))
    if net is not None:
        net.attach(ens)
    ens.reset()
    ens.set_stop_at_end(False)
    ens.set_events(1e-3, False)
    cap = 0
    for w in range(W):
        cap += R
        _progress(f"{name} warmup {w}")(ens.run(INF_T, max_attempts=cap))
    att0 = group_attempts(ens)
    time0 = [ens.state(g)[1] for g in range(len(ens.groups))]
    s0 = float(sum(t.sum() for t in time0))
    walls, st = [], None
    for k in range(K):
        cap += R
        t0 = time.perf_counter()
        st = ens.run(INF_T, max_attempts=cap)
        walls.append(time.perf_counter() - t0)
        log(f"[{name} step {k}] {walls[-1]:.3f} s; attempts {st.attempts} t_min {st.min_time:.4g} s "
            f"t_max {st.max_time:.4g} s failed {st.n_failed}")
    if st.n_failed:
        raise RuntimeError(f"{st.n_failed} instances of {name} failed")
    att1 = group_attempts(ens)
    d_att = [a1 - a0 for a0, a1 in zip(att0, att1)]
    sim_s = st.inst_time_sum - s0
    steps_total, attempts_total = int(st.steps), int(st.attempts)
    t_g, b_g = ens.time_gemm(20)
    kb, var, tt = ens.tuning()
    # per-phase split of an attempt (launched one by one, HIP events), from
    # where the timed window ended
    phases = ens.time_phases(20)
    Tslip, Tevt = args.slip_years * configs.YEAR, args.event_years * configs.YEAR
    _, ys = slip_at(ens, Tslip, f"{name} slip window")
    te = first_events(ens, Tevt, f"{name} first events")
    ens.close()
    serr = slip_error(ref["ys"], ys)
    ev = event_error(ref["te"], te)
    wall = float(np.sum(walls))
    return {"name": name, "value": sim_s / configs.YEAR / wall, "walls": walls, "wall": wall,
            "sim_years": sim_s / configs.YEAR, "d_att": d_att,
            "attempts": int(sum(a.sum() for a in d_att)), "steps_total": steps_total,
            "attempts_total": attempts_total,
            "t_window": [float(min(t.min() for t in time0)), float(st.min_time),
                         float(st.max_time)],
            "t_gemm": t_g, "b_gemm": b_g, "phases_us": {k: round(1e6 * v, 1) for k, v in phases.items()},
            "tiling": {"kb_per_split": kb, "variant": var,
                       "us_per_launch": [round(1e6 * x, 1) for x in tt]},
            "slip_error_vs_f64": serr, "first_event_time_vs_f64": ev,
            "meets_bar": meets_bar(serr, ev)}


def bench_variant(args, eng, name, net, ref, flops):
    """A secondary variant on the reference's schedule: R attempts per
    instance from the initial states, timed once; GEMM timing; slip and
    first-event errors against binary64."""
    from unicycle_amd import configs
    R = args.attempts
    ens, _ = configs.ensemble_setup(eng, _prec(name))
    if net is not None:
        net.attach(ens)
    t0 = time.perf_counter()
    st = run_attempts(ens, R, name)
    wall = time.perf_counter() - t0
    t_g, b_g = ens.time_gemm(20)
    kb, var, tt = ens.tuning()
    _, ys = slip_at(ens, args.slip_years * configs.YEAR, f"{name} slip window")
    te = first_events(ens, args.event_years * configs.YEAR, f"{name} first events")
    ens.close()
    serr = slip_error(ref["ys"], ys)
    ev = event_error(ref["te"], te)
    return {"value_from_t0": st.inst_time_sum / configs.YEAR / wall, "unit": "sim-yr/s",
            "window": f"{R} attempts per instance from the initial states",
            "us_per_lockstep_attempt": 1e6 * wall / R,
            "speed_vs_f64_per_attempt": (ref["wall"] / R) / (wall / R),
            "rejected_fraction": 1.0 - st.steps / max(1, st.attempts),
            "stage_gemm_us": 1e6 * t_g, "stage_gemm_GBps": b_g / t_g / 1e9,
            "stage_gemm_hbm_frac": b_g / t_g / HBM_PEAK,
            "stage_gemm_useful_TFLOPs": flops / t_g / 1e12,
            "tiling": {"kb_per_split": kb, "variant": var},
            "slip_error_vs_f64": serr, "slip_window_s": args.slip_years * configs.YEAR,
            "first_event_time_vs_f64": ev, "meets_bar": meets_bar(serr, ev),
            "dtype": VARIANTS[name][2]}


def target_setup(eng, precision, n_inst=256, nx=64, nz=64, seed=500, shared=True):
    """The target ensemble: n_inst instances of one nx x nz two-km thrust
    (4096 patches), with the ensemble draws; shared=True adds them as one
    geometry (slices of 64 sharing one K), False as separate groups of 64
    (each with its own copy of K)."""
    from unicycle_amd import Engine, configs
    from unicycle_amd.ensemble import Ensemble
    m = configs.thrust_fault(nx=nx, nz=nz, ellipse=None)
    d = configs.ensemble_instances(m, n_inst, seed=seed)
    ic = Engine(eng.device)
    ic.set_elastic(m.lambda_, m.mu)
    ic.set_rectangles(*m.geometry())
    y0 = []
    for i in range(n_inst):
        ic.set_friction(*configs.instance_friction(m, d, i))
        y0.append(configs.perturbed_initial_state(m, ic.initStateVector(), int(d["seeds"][i])))
    ic.close()
    y0 = np.array(y0)
    ens = Ensemble(eng, precision)
    parts = [slice(0, n_inst)] if shared else [slice(k, min(n_inst, k + 64))
                                                for k in range(0, n_inst, 64)]
    for sl in parts:
        ens.add_group(m, d["a"][sl], d["b"][sl], d["sig"][sl], d["L"][sl], y0[sl])
    ens.build()
    return ens, m


def bench_target(args, eng, net, plain_us_4096):
    """256 instances of one 4096-patch geometry in bf16+i8 (with the
    step-size network when given): the K stream shared by the four slices
    of 64 instances against four separate groups with their own K; per
    lockstep attempt and per stage GEMM; the bar against batched binary64 of
    the same ensemble (shared K)."""
    from unicycle_amd import PREC_BF16_I8, PREC_F64, configs
    R = args.target_attempts
    out = {"workload": "256 instances of one 64 x 64 thrust (4096 two-km Okada patches), "
                       f"{R} RK45 attempts per instance from the initial states",
           "instances": 256, "patches": 4096}
    Tslip, Tevt = args.slip_years * configs.YEAR, args.event_years * configs.YEAR
    res = {}
    for shared in (True, False):
        ens, m = target_setup(eng, PREC_BF16_I8, shared=shared)
        if net is not None:
            net.attach(ens)
        t0 = time.perf_counter()
        st = run_attempts(ens, R, f"target {'shared' if shared else 'separate'}")
        wall = time.perf_counter() - t0
        ph = ens.time_phases(10)
        t_gemm = float(np.mean([ph[f"gemm {k}"] for k in range(1, 7)]))
        kb, var, _ = ens.tuning()
        r = {"us_per_lockstep_attempt": 1e6 * wall / R, "stage_gemm_us": 1e6 * t_gemm,
             "families": len(ens.families), "groups": len(ens.groups),
             "tiling": {"kb_per_split": kb, "variant": var},
             "value_from_t0": st.inst_time_sum / configs.YEAR / wall,
             "attempt_phases_us": {k: round(1e6 * v, 1) for k, v in ph.items()}}
        if shared:
            _, ys = slip_at(ens, Tslip, "target slip window")
            r["te"] = first_events(ens, Tevt, "target first events")
            r["ys"] = ys
        res["shared" if shared else "separate"] = r
        ens.close()
    sh, sp = res["shared"], res["separate"]
    # K bytes of one 4096-patch geometry in bf16+i8 + row scales + 4 V slices
    mp, npad = 3 * 4096, 2 * 4096
    kbytes = 3.0 * mp * npad + 8.0 * mp + 4 * 4.0 * npad * 64
    ens, _ = target_setup(eng, PREC_F64, shared=True)
    _, ys64 = slip_at(ens, Tslip, "target f64 slip window")
    te64 = first_events(ens, Tevt, "target f64 first events")
    ens.close()
    serr = slip_error(ys64, sh.pop("ys"))
    ev = event_error(te64, sh.pop("te"))
    flops = 2.0 * mp * npad * 256
    out.update({
        "shared_k": sh, "separate_k": sp,
        "stage_gemm_speedup_shared_vs_separate": sp["stage_gemm_us"] / sh["stage_gemm_us"],
        "attempt_speedup_shared_vs_separate":
            sp["us_per_lockstep_attempt"] / sh["us_per_lockstep_attempt"],
        "shared_gemm_roofline": {
            "bytes_per_launch": kbytes, "GBps": kbytes / sh["stage_gemm_us"] / 1e3,
            "hbm_frac": kbytes / (sh["stage_gemm_us"] * 1e-6) / HBM_PEAK,
            "useful_TFLOPs": flops / sh["stage_gemm_us"] / 1e6,
            "executed_bf16_frac": 3.0 * flops / (sh["stage_gemm_us"] * 1e-6) / 2.5e15},
        "plain_fp64_us_per_attempt": plain_us_4096,
        "vs_plain_fp64_per_attempt": 256.0 * plain_us_4096 / sh["us_per_lockstep_attempt"],
        "slip_error_vs_f64": serr, "first_event_time_vs_f64": ev,
        "meets_bar": meets_bar(serr, ev),
        "learned_step_proposals": net is not None})
    return out


def plain_per_attempt(args, info):
    """The single-instance binary64 device path (streaming GEMV per
    instance): seconds per RK45 attempt on instance 0 of every group from its
    initial state."""
    from unicycle_amd import Engine, configs
    per, detail, cpu_inputs = [], [], []
    for gi, (m, d, y0) in enumerate(info):
        p = Engine(0)
        p.set_elastic(m.lambda_, m.mu)
        p.set_rectangles(*m.geometry())
        p.set_friction(*configs.instance_friction(m, d, 0))
        p.buildG()
        res = p.integrate(y0[0], INF_T, max_steps=args.plain_steps)
        per.append(res.wall_seconds / res.attempts)
        detail.append({"patches": m.n, "attempts": int(res.attempts),
                       "us_per_attempt": 1e6 * per[-1]})
        log(f"plain group {gi}: {res.attempts} attempts, {1e6 * per[-1]:.0f} us/attempt")
        if not args.no_cpu:
            cpu_inputs.append((gi, p.get_kernel(), m, d, y0[0]))
        p.close()
    return per, detail, cpu_inputs


def bench_learned(args, plain_us_per_attempt):
    """The learned step-size proposal on a fixed simulated window: the
    1024-patch group's instances 0..15 give the plain binary64 integrator's
    step histories (training data), refined by two closed-loop rounds on the
    same instances and an output offset picked on half the window; the 48
    held-out instances run as an ensemble to the window's end (clamped) with
    the classic controller and with the learned proposal, in binary64 and in
    split bf16."""
    from unicycle_amd import Engine, PREC_F64, PREC_SPLIT_BF16, configs, stepnet
    from unicycle_amd.ensemble import Ensemble
    T = args.learned_years * configs.YEAR
    nx, nz = configs.ENSEMBLE_GROUPS[0]
    m = configs.thrust_fault(nx=nx, nz=nz, ellipse=None)
    d = configs.ensemble_instances(m, 64, seed=100)   # the bench's group 0 draws
    eng = Engine(0)
    ic = Engine(0)
    ic.set_elastic(m.lambda_, m.mu)
    ic.set_rectangles(*m.geometry())
    y0 = []
    for i in range(64):
        ic.set_friction(*configs.instance_friction(m, d, i))
        y0.append(configs.perturbed_initial_state(m, ic.initStateVector(), int(d["seeds"][i])))
    y0 = np.array(y0)
    t0 = time.perf_counter()
    X, Tt = stepnet.training_data(ic, m, d, range(16), args.learned_steps, y0)
    t_data = time.perf_counter() - t0
    ic.close()
    net = stepnet.StepNet(eng, seed=0)
    net.fit_normalization(X, Tt)
    t0 = time.perf_counter()
    loss = net.train(X, Tt, epochs=40, batch=4096, lr=2e-3, seed=1)
    t_train = time.perf_counter() - t0
    rmse = float(np.sqrt(np.mean((net.forward(X) - Tt) ** 2)))
    rmse_classic = float(np.sqrt(np.mean((X[:, 6] - Tt) ** 2)))
    log(f"learned: {X.shape[0]} samples in {t_data:.1f} s, trained in {t_train:.2f} s, "
        f"rmse {rmse:.3f} (classic rule {rmse_classic:.3f})")
    # closed-loop refinement and output offset on the training instances
    t0 = time.perf_counter()
    tr = list(range(16))
    ens = Ensemble(eng, PREC_F64)
    ens.add_group(m, d["a"][tr], d["b"][tr], d["sig"][tr], d["L"][tr], y0[tr])
    ens.build()
    X, Tt = stepnet.fit_closed_loop(net, X, Tt, ens, args.learned_steps, rounds=2,
                                    progress=_progress("closed-loop"))
    offset, tried = stepnet.calibrate_offset(net, ens, window=0.5 * T, progress=_progress("offset"))
    ens.close()
    t_refine = time.perf_counter() - t0
    log(f"learned: closed-loop data {X.shape[0]} samples, offset {offset} from {tried}, "
        f"{t_refine:.1f} s")
    held = list(range(16, 64))
    runs = {}
    for name, prec, use in (("f64_classic", PREC_F64, False), ("f64_learned", PREC_F64, True),
                            ("split_bf16_learned", PREC_SPLIT_BF16, True)):
        ens = Ensemble(eng, prec)
        ens.add_group(m, d["a"][held], d["b"][held], d["sig"][held], d["L"][held], y0[held])
        ens.build()
        if use:
            net.attach(ens)
        ens.set_stop_at_end(True)
        ens.set_events(1e-3, False)
        t0 = time.perf_counter()
        st = ens.run(T, max_attempts=10**6, chunk_attempts=2000, progress=_progress(name))
        wall = time.perf_counter() - t0
        y, t, dt, steps, att = ens.state(0)
        runs[name] = {"y": y, "att": att, "steps": steps, "te": ens.first_events(0),
                      "wall": wall, "failed": st.n_failed}
        if st.n_failed:
            log(f"[{name}] failed instances, status {ens.status(0).tolist()}")
        ens.close()
    net.close()
    ref = runs["f64_classic"]
    n_h = len(held)
    plain_wall = float(np.sum(ref["att"])) * plain_us_per_attempt * 1e-6
    out = {"window_years": args.learned_years, "patches": m.n, "held_out_instances": n_h,
           "training": {"samples": int(X.shape[0]), "instances": 16,
                        "steps_per_instance": args.learned_steps, "epochs": 40,
                        "loss_first_last": [float(loss[0]), float(loss[-1])],
                        "rmse_log10_h": rmse, "rmse_classic_rule": rmse_classic,
                        "train_s": t_train, "closed_loop_rounds": 2,
                        "samples_after_closed_loop": int(X.shape[0]),
                        "output_offset_log10": offset,
                        "offset_attempts_to_half_window": {
                            str(k): (v if np.isfinite(v) else None) for k, v in tried.items()},
                        "refine_s": t_refine},
           "plain_fp64_classic_sim_yr_per_s": n_h * args.learned_years / plain_wall}
    for name, r in runs.items():
        a, s_ = float(np.sum(r["att"])), float(np.sum(r["steps"]))
        out[name] = {"sim_yr_per_s": n_h * args.learned_years / r["wall"],
                     "vs_plain_fp64": plain_wall / r["wall"],
                     "rhs_per_sim_yr_per_instance": 6.0 * a / (n_h * args.learned_years),
                     "rejected_fraction": 1.0 - s_ / a, "failed": int(r["failed"]),
                     "attempt_gain_vs_classic": float(np.sum(ref["att"])) / a,
                     "slip_error_vs_f64_classic": slip_error([ref["y"]], [r["y"]]),
                     "first_event_time_vs_f64_classic": event_error(ref["te"], r["te"])}
    return out


def cpu_baseline_ensemble(cpu_inputs, att64, sim_years, budget_s=20.0):
    """Oracle (C restatement, OpenMP) on the host cores: attempts/s of
    instance 0 of each group over a bounded sample of its run, costed over
    every instance's attempt count of the GPU run."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import orc
    from unicycle_amd import configs
    cores = host_cores()
    total, samples = 0.0, []
    for gi, K, m, d, y0 in cpu_inputs:
        sysm = orc.System(configs.instance_friction(m, d, 0), K, nthreads=cores)
        t0 = time.perf_counter()
        sysm.odefun(y0)
        t_rhs = time.perf_counter() - t0
        n_steps = max(2, int(budget_s / len(cpu_inputs) / (6.5 * t_rhs)))
        t0 = time.perf_counter()
        _, st, *_ = sysm.integrate(y0, INF_T, max_steps=n_steps, log_cap=0)
        wall = time.perf_counter() - t0
        rate = st.attempts / wall
        total += float(np.sum(att64[gi])) / rate
        samples.append(f"{m.n} patches: {st.attempts} attempts in {wall:.1f} s")
        log(f"cpu group {gi}: {rate:.2f} attempts/s")
    return {"value": sim_years / total, "unit": "sim-yr/s", "cores": cores, "kind": "port",
            "sample": "instance 0 of each group from its initial state, " + "; ".join(samples)
                      + "; costed with every instance's GPU attempt count",
            "host": host_info()}


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
    for _ in range(args.warmup):
        eng.integrate(y0, INF_T, max_steps=100)
    t_start = time.perf_counter()
    for _ in range(args.steps):
        res = eng.integrate(y0, INF_T, max_steps=args.attempts)
    t_total = time.perf_counter() - t_start
    t_k = eng.time_rhs_kernel(100)
    algo_bytes = 8.0 * m * n + 8.0 * (n + m) + 7 * 8.0 * (m // 3) * 2
    return {
        "metric": METRIC, "value": res.time / configs.YEAR / (t_total / args.steps),
        "unit": "sim-yr/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * t_total / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded thrust fault, perturbed steady state)",
        "config": {"workload": f"plain binary64, 2048-patch thrust, {args.attempts} accepted steps"},
        "roofline": {"bound": "hbm", "achieved": algo_bytes / t_k / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": algo_bytes / t_k / HBM_PEAK, "traffic": None,
                     "kernel": "ucy::gemv_post_kernel", "avg_launch_us": 1e6 * t_k},
        "detail": {"attempts": res.attempts, "us_per_attempt": 1e6 * res.wall_seconds / res.attempts},
    }


def bench_assembly(args, cpu=True):
    """Green's-function assembly (buildG, greens.f90:73-275) on the device
    for the three source types at the model sizes of the single-instance
    configurations, with the oracle's OpenMP buildG on the host cores for
    per-pair rates on smaller meshes of the same kind.  A pair is one
    (receiver, source) element pair with all its slip directions or
    eigenstrains: 2 x 3 traction entries (rectangles, triangles), 6 x 3 or
    6 x 6 (volumes).  On each model it also times one RHS with the binary64
    K and with the split-bf16 K (plain_rhs)."""
    from unicycle_amd import Engine, PREC_SPLIT_BF16, configs
    out = {}

    rhs = {}

    def rhs_time(eng, variant):
        y0 = eng.initStateVector(variant)
        eng.odefun(y0.size, 0.0, y0, np.empty_like(y0))
        return eng.time_rhs_kernel(20)

    def gpu(tag, setup, warm, fric, variant=0):
        """buildG timing, then one RHS (K v + odefun post-pass, the kernel of
        every RK45 stage) with the binary64 K and with the split-bf16 K."""
        eng = Engine(0)
        warm(eng)                         # the same kernels on a small model
        eng.buildG()
        setup(eng)
        t0 = time.perf_counter()
        eng.buildG()
        wall = time.perf_counter() - t0
        secs = eng.build_seconds()
        m, n = eng.shape
        fric(eng)
        t64 = rhs_time(eng, variant)
        eng.buildG(PREC_SPLIT_BF16)
        tsp = rhs_time(eng, variant)
        eng.close()
        rhs[tag] = {"K_rows": m, "K_cols": n,
                    "f64": {"us": 1e6 * t64, "rhs_per_s": 1.0 / t64,
                            "GBps": 8.0 * m * n / t64 / 1e9, "hbm_frac": 8.0 * m * n / t64 / HBM_PEAK},
                    "split_bf16": {"us": 1e6 * tsp, "rhs_per_s": 1.0 / tsp,
                                   "GBps": 4.0 * m * n / tsp / 1e9,
                                   "hbm_frac": 4.0 * m * n / tsp / HBM_PEAK},
                    "speedup": t64 / tsp}
        log(f"assembly {tag}: {wall:.3f} s, K {8.0 * m * n / 1e9:.1f} GB; RHS {1e6 * t64:.0f} us "
            f"binary64, {1e6 * tsp:.0f} us split bf16")
        return wall, secs, 8.0 * m * n

    # rectangles: the 2048-patch thrust (Okada traction pairs)
    rm = configs.thrust_fault()
    N = rm.n

    def s_rect(e, m=rm):
        e.set_elastic(m.lambda_, m.mu)
        e.set_rectangles(*m.geometry())
    wall, secs, kb = gpu("rectangles", s_rect, lambda e: s_rect(e, configs.thrust_fault(8, 8)),
                         lambda e: e.set_friction(*rm.friction()))
    def roof(pairs_per_s, key):
        a = pairs_per_s * ASSEMBLY_FLOP_PER_PAIR[key]
        r = {"bound": "fp64 valu", "achieved": a / 1e12, "peak": FP64_VALU_PEAK / 1e12,
             "unit": "TFLOP/s", "frac": a / FP64_VALU_PEAK,
             "flop_per_pair": ASSEMBLY_FLOP_PER_PAIR[key]}
        if key in EXECUTED_FLOP_PER_PAIR:
            x = pairs_per_s * EXECUTED_FLOP_PER_PAIR[key]
            r.update({"executed_flop_per_pair": EXECUTED_FLOP_PER_PAIR[key],
                      "executed_achieved": x / 1e12, "executed_frac": x / FP64_VALU_PEAK})
        return r

    out["okada_rectangles"] = {"elements": N, "pairs": N * N, "K_bytes": kb, "gpu_s": secs[0],
                               "gpu_pairs_per_s": N * N / secs[0],
                               "roofline": roof(N * N / secs[0], "okada")}
    # triangles: the 20,000-triangle curved megathrust (19.2 GB K)
    tm = configs.megathrust_mesh()
    T = tm.n

    def s_tri(e, m=tm):
        e.set_elastic(m.lambda_, m.mu)
        e.set_triangles(*m.geometry())
    wall, secs, kb = gpu("triangles", s_tri, lambda e: s_tri(e, configs.megathrust_mesh(8, 8)),
                         lambda e: e.set_friction(*tm.friction()))
    out["nikkhoo_triangles"] = {"elements": T, "pairs": T * T, "K_bytes": kb, "gpu_s": secs[1],
                                "gpu_pairs_per_s": T * T / secs[1],
                                "roofline": roof(T * T / secs[1], "nikkhoo")}
    # fault + strain volumes: 4096 1.5-km patches and 32 x 32 x 8 cuboids (28.2 GB K)
    fm = configs.thrust_fault(nx=64, nz=64, patch=1500.0)
    vm = configs.volume_grid()
    Nf, Nv = fm.n, vm.n

    def s_vol(e, f=fm, v=vm):
        e.set_elastic(f.lambda_, f.mu)
        e.set_rectangles(*f.geometry())
        e.set_volumes(*v.geometry())
    def f_vol(e):
        e.set_friction(*fm.friction())
        e.set_rheology(*vm.rheology())
    wall, secs, kb = gpu("fault + volumes", s_vol,
                         lambda e: s_vol(e, configs.thrust_fault(8, 8), configs.volume_grid(4, 4, 2)),
                         f_vol, variant=1)
    out["fault_and_volumes"] = {
        "patches": Nf, "volumes": Nv, "K_bytes": kb, "gpu_s_total": wall,
        "okada_pairs": Nf * (Nf + Nv), "okada_gpu_s": secs[0],
        "okada_gpu_pairs_per_s": Nf * (Nf + Nv) / secs[0],
        "volume_pairs": Nv * (Nf + Nv), "volume_gpu_s": secs[2],
        "volume_gpu_pairs_per_s": Nv * (Nf + Nv) / secs[2],
        "volume_roofline": roof(Nv * (Nf + Nv) / secs[2], "strain_volume")}
    out["plain_rhs"] = rhs
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import orc
        cores = host_cores()
        samples = {
            "okada_rectangles": (lambda: orc.Mesh(rm.lambda_, rm.mu,
                                                  rect=configs.thrust_fault(64, 32).geometry())),
            "nikkhoo_triangles": (lambda: orc.Mesh(tm.lambda_, tm.mu, tri=configs.megathrust_mesh(
                nx=24, ny=24).geometry())),
            "volumes": (lambda: orc.Mesh(fm.lambda_, fm.mu, vol=configs.volume_grid(8, 8, 4).geometry())),
        }
        for key, mk in samples.items():
            mesh = mk()
            ne = mesh.np + mesh.nvol
            t0 = time.perf_counter()
            orc.buildG(mesh, nthreads=cores)
            w = time.perf_counter() - t0
            rate = ne * ne / w
            tgt = out[key] if key in out else out["fault_and_volumes"]
            tgt["cpu_pairs_per_s" if key in out else "volume_cpu_pairs_per_s"] = rate
            tgt["cpu_sample" if key in out else "volume_cpu_sample"] = (
                f"oracle buildG, {ne} elements ({ne * ne} pairs) in {w:.2f} s on {cores} cores")
            log(f"assembly cpu {key}: {rate:.3g} pairs/s")
    return out


def bench_gemm(args):
    """The all-groups stage GEMM of one precision alone, launched directly
    (no graph): autotuning launches, then 20 timed launches.  For profiler
    and counter runs."""
    from unicycle_amd import Engine, PREC_BF16_I8, PREC_F64, PREC_SPLIT_BF16, configs
    prec = {"f64": PREC_F64, "split_bf16": PREC_SPLIT_BF16,
            "bf16_i8": PREC_BF16_I8}[args.precision]
    eng = Engine(0)
    ens, info = configs.ensemble_setup(eng, prec)
    t, b = ens.time_gemm(20)
    kb, var, tt = ens.tuning()
    flops = sum(2.0 * 3 * m.n * 2 * m.n * 64 for m, d, y0 in info)
    ens.close()
    return {"workload": f"stage GEMM {args.precision}", "avg_launch_us": 1e6 * t,
            "bytes_per_launch": b, "GBps": b / t / 1e9, "useful_TFLOPs": flops / t / 1e12,
            "tuning": {"kb_per_split": kb, "variant": var,
                       "us_per_launch": [round(1e6 * x, 1) for x in tt]}}


def bench_target_gemm(args):
    """The stage GEMM of the 256-instance, 4096-patch target alone (one K
    family of four slices, bf16+i8), launched directly (no graph):
    autotuning launches, then 20 timed launches.  For profiler and counter
    runs."""
    from unicycle_amd import Engine, PREC_BF16_I8
    eng = Engine(0)
    ens, m = target_setup(eng, PREC_BF16_I8, shared=True)
    t, b = ens.time_gemm(20)
    kb, var, tt = ens.tuning()
    flops = 2.0 * 3 * (3 * m.n) * (2 * m.n) * 256
    ens.close()
    return {"workload": "stage GEMM, 256 instances of one 4096-patch geometry (bf16+i8, one "
                        "K family of 4 slices)", "avg_launch_us": 1e6 * t,
            "bytes_per_launch": b, "GBps": b / t / 1e9,
            "executed_TFLOPs": flops / t / 1e12, "executed_bf16_frac": flops / t / 2.5e15,
            "tuning": {"kb_per_split": kb, "variant": var,
                       "us_per_launch": [round(1e6 * x, 1) for x in tt]}}


def bench_rhs(args):
    """The plain path's fused matvec + odefun post-pass kernel alone on the
    2048-patch thrust, launched directly (no graph), 100 timed launches."""
    from unicycle_amd import Engine, configs
    model = configs.thrust_fault(nx=64, nz=32)
    eng = Engine(0)
    eng.set_elastic(model.lambda_, model.mu)
    eng.set_rectangles(*model.geometry())
    eng.set_friction(*model.friction())
    eng.buildG()
    m, n = eng.shape
    y0 = configs.perturbed_initial_state(model, eng.initStateVector(), seed=1)
    eng.odefun(y0.size, 0.0, y0, np.empty_like(y0))
    t_k = eng.time_rhs_kernel(100)
    algo_bytes = 8.0 * m * n + 8.0 * (n + m) + 7 * 8.0 * (m // 3) * 2
    return {"workload": "plain RHS kernel", "avg_launch_us": 1e6 * t_k,
            "bytes_per_launch": algo_bytes, "GBps": algo_bytes / t_k / 1e9}


def bench_attempt(args):
    """The kernels of an RK45 attempt of the headline ensemble (bf16+i8 K,
    all groups), launched one by one rather than from the captured graph:
    R attempts from the initial states, then --steps x 20 attempts with HIP
    events between the kernels.  For profiler runs of the attempt's kernels."""
    from unicycle_amd import PREC_BF16_I8, configs
    from unicycle_amd import Engine
    eng = Engine(0)
    ens, info = configs.ensemble_setup(eng, PREC_BF16_I8)
    run_attempts(ens, args.attempts, "attempt workload")
    ph = {}
    for _ in range(max(1, args.steps)):
        for k, v in ens.time_phases(20).items():
            ph[k] = ph.get(k, 0.0) + v / max(1, args.steps)
    ens.close()
    eng.close()
    return {"workload": "attempt kernels, bf16+i8 ensemble (4 x 64 instances)",
            "phases_us": {k: round(1e6 * v, 1) for k, v in ph.items()},
            "attempt_us": round(1e6 * sum(ph.values()), 1)}


def bench_copy(args):
    """Calibration of the HBM byte counters: a device copy of known size."""
    import torch
    n = 1 << 28                      # 2 GiB of binary64
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    return {"workload": "copy", "bytes_read_per_launch": 8 * n, "bytes_written_per_launch": 8 * n}


def bench_ensemble(args):
    """The headline line: batched binary64 once (the precision reference),
    the step-size network, the headline variant's timed steps, the plain
    binary64 device path and the oracle on the host cores per attempt; then,
    while the wall budget lasts, the other variants, the learned-proposal
    comparison and the assembly figures."""
    from unicycle_amd import Engine, configs
    budget = Budget(args.budget)
    R, W, K = args.attempts, args.warmup, args.steps
    out = {"metric": METRIC, "value": None, "unit": "sim-yr/s", "n_gpus": 1, "steps": K,
           "warmup": W, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": None,
           "data": "synthetic (seeded friction draws and initial slip-rate perturbations)",
           "config": {"workload": "batched ensemble: 4 geometry groups x 64 instances of "
                                  "1024/2048/4096/8192 2-km Okada patches (30 deg thrust); "
                                  f"a step = {R} RK45 attempts per instance continuing one run "
                                  f"from the seeded initial states ({W} untimed, then {K} timed "
                                  "steps); eps 1e-6",
                      "attempts_per_step": R}}
    eng = Engine(0)
    shapes = []
    ref = section(out, "batched_f64", budget, 0.0, lambda: bench_reference(args, eng, shapes))
    if ref is None:
        return out
    flops = sum(2.0 * m * n * 64 for m, n, _ in shapes)
    out["config"].update({"instances": sum(s[2] for s in shapes),
                          "groups": [[m, n] for m, n, _ in shapes]})
    name, net, net_info = args.headline, None, None
    if VARIANTS[name][1]:
        r = section(out, "step_net_training", budget, 0.0, lambda: train_step_net(eng, args))
        if r is None:
            name = "bf16_i8"
        else:
            net, net_info = r
    head = section(out, "headline", budget, 0.0,
                   lambda: bench_headline(args, eng, name, net, ref))
    pl = section(out, "plain_fp64", budget, 0.0, lambda: plain_per_attempt(args, ref["info"]))

    f64_costed = None
    if head is not None:
        # batched binary64 costed per lockstep attempt over the headline's window
        f64_costed = head["sim_years"] / (K * ref["wall"])
        out["config"]["headline_variant"] = name
        out["config"]["gemm_tiling"] = head["tiling"]
        out["config"]["window_s"] = {"t_min_start": head["t_window"][0],
                                     "t_min_end": head["t_window"][1],
                                     "t_max_end": head["t_window"][2]}
        hd = {"value": head["value"], "unit": "sim-yr/s", "meets_bar": head["meets_bar"],
              "slip_error_vs_f64": head["slip_error_vs_f64"],
              "slip_window_s": args.slip_years * configs.YEAR,
              "first_event_time_vs_f64": head["first_event_time_vs_f64"],
              "event_horizon_s": args.event_years * configs.YEAR,
              "step_walls_s": [round(w, 4) for w in head["walls"]],
              "simulated_years_timed": head["sim_years"], "attempts_timed": head["attempts"],
              "rhs_per_s": 6.0 * head["attempts"] / head["wall"],
              "attempt_phases_us": head["phases_us"],
              "us_per_lockstep_attempt": 1e6 * head["wall"] / (K * R),
              "rejected_fraction_since_t0": 1.0 - head["steps_total"] / max(
                  1, head["attempts_total"]),
              "dtype": VARIANTS[name][2]}
        if net_info is not None:
            hd["training"] = net_info
        out[name] = hd
        if head["meets_bar"]:
            out["value"] = head["value"]
            out["ms_per_step"] = 1e3 * float(np.mean(head["walls"]))
            out["dtype"] = "bf16+i8 K, fp32 accumulate, f64 state"
            if name == "split_bf16":
                out["dtype"] = "split-bf16 K, fp32 accumulate, f64 state"
            # the stage GEMM as it runs inside the attempts: mean of the six
            # GEMM phases (HIP events between the attempt's kernels)
            t_in = 1e-6 * float(np.mean([head["phases_us"][f"gemm {s}"] for s in range(1, 7)]))
            gb = head["b_gemm"] / t_in / 1e9
            fmt = "bf16_i8" if name.startswith("bf16_i8") else "split_bf16"
            out["roofline"] = {
                "bound": "hbm", "achieved": gb, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": gb * 1e9 / HBM_PEAK, "traffic": MEASURED_TRAFFIC.get(fmt),
                "kernel": f"ucy::gemm_bf16_kernel<{head['tiling']['variant']}, "
                          f"{str(fmt == 'bf16_i8').lower()}> (stage GEMM, all groups)",
                "avg_launch_us": 1e6 * t_in, "bytes_per_launch": head["b_gemm"],
                "useful_TFLOPs": flops / t_in / 1e12,
                "timing": "HIP events around each kernel of 20 attempts launched one by one "
                          "on the library stream, right after the timed steps (mean of the 6 "
                          "stage GEMMs)",
                "back_to_back_us": 1e6 * head["t_gemm"]}
        else:
            out.setdefault("errors", {})["headline"] = (
                f"{name} misses the bar (slip {head['slip_error_vs_f64']}, events "
                f"{head['first_event_time_vs_f64']}); value is batched binary64")
            out["value"] = f64_costed
            out["ms_per_step"] = 1e3 * ref["wall"]
            out["dtype"] = "f64"
            out["config"]["headline_variant"] = "f64"
    out["batched_f64"] = {
        "value_from_t0": ref["st"].inst_time_sum / configs.YEAR / ref["wall"],
        "window_from_t0": f"{R} attempts per instance from the initial states, timed once",
        "us_per_lockstep_attempt": 1e6 * ref["wall"] / R,
        "value_on_headline_window": f64_costed, "unit": "sim-yr/s",
        "setup_s": ref["setup_s"], "kb_per_split": ref["kb"]}
    out["f64_gemm_roofline"] = {
        "bound": "mfma", "achieved": flops / ref["t_gemm"] / 1e12,
        "peak": FP64_MFMA_PEAK / 1e12, "unit": "TFLOP/s", "frac": flops / ref["t_gemm"] / FP64_MFMA_PEAK,
        "avg_launch_us": 1e6 * ref["t_gemm"], "hbm_GBps": ref["gemm_bytes"] / ref["t_gemm"] / 1e9}
    cpu_inputs = []
    if pl is not None and head is not None:
        per, detail, cpu_inputs = pl
        plain_wall = sum(p * float(a.sum()) for p, a in zip(per, head["d_att"]))
        value_plain = head["sim_years"] / plain_wall
        out["plain_fp64"] = {"value": value_plain, "unit": "sim-yr/s",
                             "method": "single-instance binary64 device path (streaming GEMV), "
                                       "seconds per RK45 attempt of instance 0 of each group "
                                       "times every instance's attempts in the headline window",
                             "groups": detail}
        out["vs_plain_fp64"] = (out["value"] or 0.0) / value_plain
        out["batched_f64"]["vs_plain_fp64"] = f64_costed / value_plain
    if not args.no_cpu and cpu_inputs and head is not None:
        cb = section(out, "cpu_baseline", budget, 0.0,
                     lambda: cpu_baseline_ensemble(cpu_inputs, head["d_att"], head["sim_years"]))
        if cb is not None:
            out["cpu_baseline"] = cb
    # ---- optional sections under the wall budget
    if not args.no_target and pl is not None:
        r = section(out, "target_ensemble", budget, 90.0,
                    lambda: bench_target(args, eng, net, 1e6 * pl[0][2]))
        if r is not None:
            out["target_ensemble"] = r
    bar_s = out["section_seconds"].get("headline", 60.0) - (head["wall"] if head else 0.0)
    per_variant = max(20.0, bar_s) + 4.0 * R * 2.5e-3 + 10.0
    extra = ["bf16_i8", "split_bf16", "mixed"] if not args.no_variants else []
    for v in extra:
        if v == name:
            continue
        r = section(out, v, budget, per_variant * (2.0 if v == "mixed" else 1.0),
                    lambda v=v: bench_variant(args, eng, v, None, ref, flops))
        if r is not None:
            out[v] = r
    if net is not None:
        net.close()
    if not args.no_learned and pl is not None:
        r = section(out, "learned_step_proposal", budget, 150.0,
                    lambda: bench_learned(args, 1e6 * pl[0][0]))
        if r is not None:
            out["learned_step_proposal"] = r
    if not args.no_assembly:
        r = section(out, "assembly", budget, 100.0,
                    lambda: bench_assembly(args, cpu=not args.no_cpu))
        if r is not None:
            out["assembly"] = r
    out["section_seconds"]["total"] = round(budget.elapsed(), 1)
    eng.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5, help="timed steps of the headline variant")
    ap.add_argument("--warmup", type=int, default=2, help="untimed steps before them")
    ap.add_argument("--attempts", type=int, default=1000,
                    help="RK45 attempts per instance in one step")
    ap.add_argument("--headline", default="bf16_i8_learned", choices=sorted(VARIANTS))
    ap.add_argument("--budget", type=float, default=420.0,
                    help="wall seconds after which optional sections are skipped")
    ap.add_argument("--slip-years", type=float, default=0.05,
                    help="pre-event time of the slip-error comparison")
    ap.add_argument("--event-years", type=float, default=0.5,
                    help="horizon of the first-event comparison")
    ap.add_argument("--plain-steps", type=int, default=400,
                    help="accepted steps of the plain path's per-attempt timing")
    ap.add_argument("--workload", default="ensemble",
                    choices=["ensemble", "plain", "gemm", "target_gemm", "rhs", "copy",
                             "assembly", "attempt"])
    ap.add_argument("--precision", default="split_bf16", choices=["f64", "split_bf16", "bf16_i8"],
                    help="stage-GEMM precision of --workload gemm")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true",
                    help="no learned-proposal section (the headline keeps its network)")
    ap.add_argument("--no-variants", action="store_true")
    ap.add_argument("--no-assembly", action="store_true")
    ap.add_argument("--no-target", action="store_true",
                    help="no 256 x 4096-patch shared-K section")
    ap.add_argument("--target-attempts", type=int, default=300,
                    help="RK45 attempts per instance of the shared-K section")
    ap.add_argument("--learned-years", type=float, default=0.25,
                    help="simulated window of the learned-proposal comparison")
    ap.add_argument("--learned-steps", type=int, default=3000,
                    help="plain-integrator steps per training instance")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the engine runs on a single MI355X and has no "
              f"multi-GPU path (--gpus {args.gpus} requested)", file=sys.stderr)
        sys.exit(2)
    if args.workload != "ensemble":
        fn = {"plain": bench_plain, "gemm": bench_gemm, "target_gemm": bench_target_gemm,
              "rhs": bench_rhs,
              "copy": bench_copy, "attempt": bench_attempt,
              "assembly": lambda a: bench_assembly(a, cpu=not a.no_cpu)}[args.workload]
        print(json.dumps(fn(args)), flush=True)
        return
    out = bench_ensemble(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
