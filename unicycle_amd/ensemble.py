"""Batched ensembles of independent fault instances (no counterpart in the
reference; each instance runs the reference's per-instance algorithm).

Instances are grouped by geometry; a group shares one kernel K.  Precisions:
  PREC_F64         binary64 K on v_mfma_f64_16x16x4f64 for every matvec;
  PREC_SPLIT_BF16  split bf16 hi+lo K with exact row scales on
                   v_mfma_f32_32x32x16_bf16; the stage matvecs are increments
                   t_i = t_0 + K (v_i - v_0) over the step start, so the
                   embedded error estimate (whose weights sum to zero) does
                   not see the constant part of the rounding;
  PREC_MIXED       as PREC_SPLIT_BF16 with t_0 = K v_0 in binary64;
  PREC_BF16_I8     as PREC_SPLIT_BF16 with K in 3 bytes per entry: the bf16
                   hi and an 8-bit code of the residual in steps of 1/128 of
                   hi's ulp, decoded to the bf16 lo operand in registers.
State, time, step control and error norms stay binary64 per instance.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
from dataclasses import dataclass

import numpy as np

from . import _lib
from .engine import (Engine, UcyError, PREC_F64, PREC_SPLIT_BF16, PREC_MIXED,  # noqa: F401
                     PREC_BF16_I8, _p)


@dataclass
class EnsembleStats:
    inst_time_sum: float
    min_time: float
    max_time: float
    steps: int
    attempts: int
    replays: int
    n_inst: int
    n_failed: int
    wall_seconds: float


class Ensemble:
    def __init__(self, engine: Engine, precision: int = PREC_SPLIT_BF16):
        self.eng = engine
        self._L = engine._L
        self._h = self._L.ucy_ensemble_create(engine._h, precision)
        if not self._h:
            raise UcyError(-1, "cannot create ensemble")
        self.groups = []
        self.families = []     # group indices sharing one K, per add_group call
        self._log_cap = {}
        engine._ensembles.append(self)

    def close(self):
        # the context destroys its ensembles with it; only a live context's
        # ensemble handle is still valid
        if getattr(self, "_h", None):
            if getattr(self.eng, "_h", None):
                self._L.ucy_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, st):
        self.eng._chk(st)

    def add_group(self, model, a, b, sig, L, y0):
        """model: configs.FaultModel geometry and the shared Vpl/rake/damping;
        a, b, sig, L: (n_inst, N) per-instance friction; y0: (n_inst, 7N).
        Up to 256 instances: beyond 64 the library stores the geometry as
        consecutive groups of 64 instances (slices, the last one partly
        filled) that share one K, and each slice appears here as a group of
        its own.  Returns the index of the first slice."""
        a = np.ascontiguousarray(a, dtype=np.float64)
        n_inst, N = a.shape
        rep = lambda v: np.ascontiguousarray(np.broadcast_to(v, (n_inst, N)), dtype=np.float64)
        arrs = [rep(model.Vpl), rep(model.rake), a, rep(b), rep(sig), rep(L), rep(model.damping),
                np.ascontiguousarray(y0, dtype=np.float64).reshape(n_inst, 7 * N)]
        geo = [np.ascontiguousarray(g, dtype=np.float64) for g in model.geometry()]
        st = self._L.ucy_ensemble_add_group(self._h, N, *[_p(g) for g in geo],
                                            float(model.lambda_), float(model.mu), n_inst,
                                            *[_p(v) for v in arrs])
        if st < 0:
            self._chk(st)
        for k in range(0, n_inst, 64):
            self.groups.append((N, min(64, n_inst - k)))
        self.families.append(list(range(st, len(self.groups))))
        return st

    def build(self):
        self._chk(self._L.ucy_ensemble_build(self._h))

    def reset(self, dt0: float = 1e-3):
        self._chk(self._L.ucy_ensemble_reset(self._h, float(dt0)))

    def run(self, t_end: float, max_attempts: int = 10**9, epsilon: float = 1e-6,
            max_step: float = 1.7976931348623158e308, chunk_attempts: int | None = None,
            progress=None) -> EnsembleStats:
        """Advance to t_end (ratestate.f90:307-328 per instance).  With
        chunk_attempts the run is issued in slices of that many attempts per
        instance, calling progress(stats) after each slice; wall_seconds then
        sums the slices."""
        if not chunk_attempts:
            return self._run(t_end, max_attempts, epsilon, max_step)
        wall, cap = 0.0, 0
        while True:
            cap = min(max_attempts, cap + int(chunk_attempts))
            st = self._run(t_end, cap, epsilon, max_step)
            wall += st.wall_seconds
            st = dataclasses.replace(st, wall_seconds=wall)
            if progress is not None:
                progress(st)
            # replays == 0: every instance is at t_end or has failed
            if st.min_time >= t_end or cap >= max_attempts or st.replays == 0:
                return st

    def _run(self, t_end, max_attempts, epsilon, max_step) -> EnsembleStats:
        st = _lib.EnsembleStats()
        self._chk(self._L.ucy_ensemble_run(self._h, float(t_end), int(max_attempts),
                                           float(epsilon), float(max_step), C.byref(st)))
        return EnsembleStats(st.inst_time_sum, st.min_time, st.max_time, st.steps, st.attempts,
                             st.replays, st.n_inst, st.n_failed, st.wall_seconds)

    def set_stop_at_end(self, on: bool = True):
        """Clamp each instance's last step to land exactly on t_end (a
        measurement option; the reference's loop steps past the interval)."""
        self._chk(self._L.ucy_ensemble_set_stop_at_end(self._h, int(bool(on))))

    def set_events(self, vmax_threshold: float = 1e-3, stop_at_event: bool = False):
        """Record each instance's first event (step-start vMax above the
        threshold); optionally stop the instance there."""
        self._chk(self._L.ucy_ensemble_set_events(self._h, float(vmax_threshold),
                                                  int(bool(stop_at_event))))

    def first_events(self, group: int) -> np.ndarray:
        _, n_inst = self.groups[group]
        t = np.empty(n_inst)
        self._chk(self._L.ucy_ensemble_first_events(self._h, group, _p(t)))
        return t

    def status(self, group: int) -> np.ndarray:
        """0 running/finished, 1 step-size underflow, 2 non-finite error norm."""
        _, n_inst = self.groups[group]
        st = np.empty(n_inst, dtype=np.int32)
        self._chk(self._L.ucy_ensemble_status(self._h, group,
                                              st.ctypes.data_as(C.POINTER(C.c_int))))
        return st

    def set_step_log(self, cap: int, group: int | None = None):
        """Log up to cap accepted steps per instance from the last reset
        (cap = 0: off), on one group or all."""
        for g in ([group] if group is not None else range(len(self.groups))):
            self._chk(self._L.ucy_ensemble_set_step_log(self._h, g, int(cap)))
            self._log_cap[g] = int(cap)

    def step_log(self, group: int) -> list[dict]:
        """Per instance: arrays t (after the step), dt, vmax (step start),
        errmax (/ epsilon) and rejects of the logged steps."""
        _, n_inst = self.groups[group]
        cap = self._log_cap.get(group, 0)
        arrs = [np.zeros((n_inst, cap)) for _ in range(5)]
        cnt = np.zeros(n_inst, dtype=np.int64)
        P = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
        self._chk(self._L.ucy_ensemble_get_step_log(
            self._h, group, cap, *[P(a) for a in arrs],
            cnt.ctypes.data_as(C.POINTER(C.c_longlong))))
        names = ("t", "dt", "vmax", "errmax", "rejects")
        return [{k: a[i, :cnt[i]].copy() for k, a in zip(names, arrs)} for i in range(n_inst)]

    def state(self, group: int):
        N, n_inst = self.groups[group]
        y = np.empty((n_inst, 7 * N))
        t = np.empty(n_inst)
        dt = np.empty(n_inst)
        steps = np.empty(n_inst, dtype=np.int64)
        att = np.empty(n_inst, dtype=np.int64)
        self._chk(self._L.ucy_ensemble_get_state(
            self._h, group, _p(y), _p(t), _p(dt),
            steps.ctypes.data_as(C.POINTER(C.c_int64)), att.ctypes.data_as(C.POINTER(C.c_int64))))
        return y, t, dt, steps, att

    def matvec(self, group: int, V) -> np.ndarray:
        N, n_inst = self.groups[group]
        V = np.ascontiguousarray(V, dtype=np.float64).reshape(2 * N, n_inst)
        T = np.empty((3 * N, n_inst))
        self._chk(self._L.ucy_ensemble_matvec(self._h, group, _p(V), _p(T)))
        return T

    def tuning(self):
        """(k-blocks per split, bf16 GEMM variant, seconds per launch of the
        candidates: k-blocks 32, 48, 64, 80, 96, 104, 128, 160, 192, 256, 512
        of variant 0, then 1, ...).  Variants: 0 = 64 K rows per wave, 1 = 64 rows with loads
        issued a k-block ahead, 2 = 128 rows, ahead, 3 = 128 rows with V
        staged in LDS in chunks of 4 k-blocks, 4 = 64 rows, ahead, V staged
        in chunks of 8, 5 = 128 rows, 6 = 64 rows and 7 = 128 rows with
        every load a direct global->LDS load (K per wave, V per block), 4 or
        3 k-blocks deep, 8 = 128 rows with K per wave by direct
        global->LDS loads and V in registers, 3 k-blocks deep, no barrier
        (6-8: bf16+i8 only), 9 = 128 rows in two register sets used in
        turn, 10 = 128 rows per block of 4 waves, one K-family slice per
        wave, each K tile loaded and decoded once per block into LDS."""
        kb, var = C.c_int(), C.c_int()
        t = np.full(_lib.GEMM_TUNE_ENTRIES, -1.0)
        self._chk(self._L.ucy_ensemble_tuning(self._h, C.byref(kb), C.byref(var), _p(t)))
        return kb.value, var.value, t

    def set_tiling(self, kb_per_split: int, variant: int = 0):
        """Override the autotuner's GEMM tiling."""
        self._chk(self._L.ucy_ensemble_set_tiling(self._h, int(kb_per_split), int(variant)))

    def time_gemm(self, iters: int = 20):
        s, b = C.c_double(), C.c_double()
        self._chk(self._L.ucy_ensemble_time_gemm(self._h, int(iters), C.byref(s), C.byref(b)))
        return s.value, b.value

    def time_phases(self, attempts: int = 20) -> dict:
        """Seconds per attempt of each phase of an RK45 attempt (kernels
        launched one by one, HIP events between them).  Advances every
        instance by `attempts` attempts."""
        t = np.zeros(len(_lib.ENS_PHASES))
        self._chk(self._L.ucy_ensemble_time_phases(self._h, int(attempts), _p(t)))
        return {k: float(v) / attempts for k, v in zip(_lib.ENS_PHASES, t)}
