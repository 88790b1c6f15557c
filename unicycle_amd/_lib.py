"""ctypes binding of libucy.so (the C ABI declared in include/ucy.h).

The HIP library is the only compute path: if it is missing, importing this
module raises, there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("UCY_LIB", os.path.join(_HERE, "libucy.so"))

_d = C.c_double
_i = C.c_int
_i64 = C.c_int64
_pd = C.POINTER(C.c_double)
_pi = C.POINTER(C.c_int)
_vp = C.c_void_p
_pf = C.POINTER(C.c_float)


class EnsembleStats(C.Structure):
    _fields_ = [
        ("inst_time_sum", C.c_double),
        ("min_time", C.c_double),
        ("max_time", C.c_double),
        ("steps", C.c_int64),
        ("attempts", C.c_int64),
        ("replays", C.c_int64),
        ("n_inst", C.c_int64),
        ("n_failed", C.c_int64),
        ("wall_seconds", C.c_double),
    ]


class RunStats(C.Structure):
    _fields_ = [
        ("time", C.c_double),
        ("dt_next", C.c_double),
        ("steps", C.c_int64),
        ("attempts", C.c_int64),
        ("rhs_evals", C.c_int64),
        ("vmax", C.c_double),
        ("wall_seconds", C.c_double),
    ]


class StepLog(C.Structure):
    _fields_ = [
        ("t", C.POINTER(C.c_double)),
        ("dt", C.POINTER(C.c_double)),
        ("vmax", C.POINTER(C.c_double)),
        ("errmax", C.POINTER(C.c_double)),
        ("rejects", C.POINTER(C.c_double)),
        ("cap", C.c_int64),
        ("n_obs", C.c_int64),
        ("obs_index", C.POINTER(C.c_int64)),
        ("obs_y", C.POINTER(C.c_double)),
        ("obs_dydt", C.POINTER(C.c_double)),
        ("disp", C.POINTER(C.c_double)),
        ("disp_f", C.POINTER(C.c_double)),
        ("disp_l", C.POINTER(C.c_double)),
        ("emax", C.POINTER(C.c_double)),
    ]


_SIGS = {
    "ucy_create": (_vp, [_i, _i]),
    "ucy_destroy": (None, [_vp]),
    "ucy_last_error": (C.c_char_p, [_vp]),
    "ucy_version": (C.c_char_p, []),
    "ucy_okada92_traction_kernels": (_i, [_vp, _pd, _pd, _pd, _pd, _i, _pd, _pd, _pd, _pd, _pd,
                                          _d, _d, _d, _d, _d, _pd, _pd, _pd]),
    "ucy_okada92_stress_kernels": (_i, [_vp, _pd, _pd, _pd, _pd, _i, _pd, _pd, _pd, _pd, _pd,
                                        _d, _d, _d, _d, _d, _pd, _pd, _pd, _pd, _pd, _pd]),
    "ucy_nikkhoo15_traction_kernels": (_i, [_vp, _pd, _pd, _pd, _pd, _i, _pi, _pi, _pi, _i, _pd,
                                            _d, _d, _d, _d, _d, _pd, _pd, _pd]),
    "ucy_nikkhoo15_stress_kernels": (_i, [_vp, _pd, _pd, _pd, _pd, _i, _pi, _pi, _pi, _i, _pd,
                                          _d, _d, _d, _d, _d, _pd, _pd, _pd, _pd, _pd, _pd]),
    "ucy_straininhalfspace": (_i, [_vp, _i, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _d, _d, _d,
                                   _d, _d]),
    "ucy_strainvolume_traction_kernels": (_i, [_vp, _pd, _pd, _pd, _pd, _i, _pd, _pd, _pd, _pd,
                                               _pd, _d, _d, _d, _d, _d, _d, _d, _d, _pd, _pd,
                                               _pd]),
    "ucy_strainvolume_stress_kernels": (_i, [_vp, _pd, _pd, _pd, _pd, _i, _pd, _pd, _pd, _pd,
                                             _pd, _d, _d, _d, _d, _d, _d, _d, _d, _pd, _pd, _pd,
                                             _pd, _pd, _pd]),
    "ucy_reference_system_okada92": (_i, [_i, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pd]),
    "ucy_reference_system_nikkhoo15": (_i, [_i, _pd, _i, _pi, _pi, _pi, _pd, _pd, _pd, _pd]),
    "ucy_reference_system_strainvolume": (_i, [_i, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pd]),
    "ucy_set_elastic": (_i, [_vp, _d, _d]),
    "ucy_set_rectangles": (_i, [_vp, _i, _pd, _pd, _pd, _pd, _pd]),
    "ucy_set_triangles": (_i, [_vp, _i, _pi, _pi, _pi, _i, _pd]),
    "ucy_set_antiplane": (_i, [_vp, _d, _i, _pd, _pd, _pd]),
    "ucy_set_antiplane_volumes": (_i, [_vp, _i, _pd, _pd, _pd, _pd]),
    "ucy_set_planestrain": (_i, [_vp, _d, _d, _i, _pd, _pd, _pd]),
    "ucy_set_initial_traction": (_i, [_vp, _i, _pd]),
    "ucy_set_planestrain_volumes": (_i, [_vp, _i, _pd, _pd, _pd, _pd]),
    "ucy_set_initial_volume_stress": (_i, [_vp, _i, _pd]),
    "ucy_planestrain_traction_kernels": (_i, [_vp, _pd, _pd, _pd, _pd, _i, _pd, _pd, _pd, _d, _d,
                                              _pd, _pd]),
    "ucy_antiplane_traction_kernels": (_i, [_vp, _pd, _pd, _pd, _pd, _i, _pd, _pd, _pd, _d, _pd]),
    "ucy_set_volumes": (_i, [_vp, _i, _pd, _pd, _pd, _pd, _pd]),
    "ucy_set_friction": (_i, [_vp, _i, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pd]),
    "ucy_set_rheology": (_i, [_vp, _i, _pd, _pd, _pd, _pd, _pd, _pd, _pd]),
    "ucy_build_kernel": (_i, [_vp, _i]),
    "ucy_kernel_shape": (_i, [_vp, C.POINTER(_i64), C.POINTER(_i64)]),
    "ucy_build_seconds": (_i, [_vp, _pd, _pd, _pd]),
    "ucy_get_kernel": (_i, [_vp, _pd]),
    "ucy_set_kernel": (_i, [_vp, _pd, _i64, _i64]),
    "ucy_kernel_dev": (_vp, [_vp]),
    "ucy_build_displacement": (_i, [_vp, _i, _pd]),
    "ucy_displacement_shape": (_i, [_vp, C.POINTER(_i64), C.POINTER(_i64)]),
    "ucy_get_displacement_kernel": (_i, [_vp, _pd, _pd, _pd]),
    "ucy_displacement": (_i, [_vp, _pd, _pd, _pd, _pd]),
    "ucy_displacement_dev": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "ucy_matvec_f64": (_i, [_vp, _pd, _pd]),
    "ucy_matvec_f64_dev": (_i, [_vp, _vp, _vp]),
    "ucy_matvec_reduced": (_i, [_vp, _pd, _pd]),
    "ucy_init_state": (_i, [_vp, _i, _pd]),
    "ucy_odefun": (_i, [_vp, _i, _d, _pd, _pd, _pd]),
    "ucy_odefun_dev_static": (_i, [_vp, _i, _vp, _vp, _vp, _vp]),
    "ucy_rkck": (_i, [_vp, _i, _d, _pd, _pd, _d, _pd, _pd]),
    "ucy_rkqs": (_i, [_vp, _i, _pd, _pd, _pd, _pd, _d, _pd, _pd, _d, _d]),
    "ucy_integrate": (_i, [_vp, _pd, _d, _d, _i64, _d, _d, _pd, _pd, _pd, _i64,
                           C.POINTER(RunStats)]),
    "ucy_integrate_ex": (_i, [_vp, _pd, _d, _d, _d, _i64, _d, _d, C.POINTER(StepLog),
                              C.POINTER(RunStats)]),
    "ucy_time_rhs_kernel": (_i, [_vp, _i, _pd]),
    "ucy_mlp_create": (_vp, [_vp, C.c_uint64]),
    "ucy_mlp_destroy": (None, [_vp]),
    "ucy_mlp_set_params": (_i, [_vp, _pf]),
    "ucy_mlp_get_params": (_i, [_vp, _pf]),
    "ucy_mlp_set_normalization": (_i, [_vp, _pf, _pf, C.c_float, C.c_float]),
    "ucy_mlp_forward": (_i, [_vp, _i64, _pf, _pf]),
    "ucy_mlp_loss_grad": (_i, [_vp, _i64, _pf, _pf, _pd, _pf]),
    "ucy_mlp_train": (_i, [_vp, _i64, _pf, _pf, _i, _i, C.c_float, C.c_uint64, _pd]),
    "ucy_ensemble_set_mlp": (_i, [_vp, _vp]),
    "ucy_ensemble_create": (_vp, [_vp, _i]),
    "ucy_ensemble_destroy": (None, [_vp]),
    "ucy_ensemble_add_group": (_i, [_vp, _i, _pd, _pd, _pd, _pd, _pd, _d, _d, _i, _pd, _pd, _pd,
                                    _pd, _pd, _pd, _pd, _pd]),
    "ucy_ensemble_build": (_i, [_vp]),
    "ucy_ensemble_reset": (_i, [_vp, _d]),
    "ucy_ensemble_run": (_i, [_vp, _d, _i64, _d, _d, C.POINTER(EnsembleStats)]),
    "ucy_ensemble_get_state": (_i, [_vp, _i, _pd, _pd, _pd, C.POINTER(_i64), C.POINTER(_i64)]),
    "ucy_ensemble_set_stop_at_end": (_i, [_vp, _i]),
    "ucy_ensemble_set_events": (_i, [_vp, C.c_double, _i]),
    "ucy_ensemble_first_events": (_i, [_vp, _i, _pd]),
    "ucy_ensemble_status": (_i, [_vp, _i, _pi]),
    "ucy_ensemble_set_step_log": (_i, [_vp, _i, C.c_longlong]),
    "ucy_ensemble_get_step_log": (_i, [_vp, _i, C.c_longlong, _pd, _pd, _pd, _pd, _pd,
                                       C.POINTER(C.c_longlong)]),
    "ucy_ensemble_matvec": (_i, [_vp, _i, _pd, _pd]),
    "ucy_ensemble_tuning": (_i, [_vp, C.POINTER(_i), C.POINTER(_i), _pd]),
    "ucy_ensemble_set_tiling": (_i, [_vp, _i, _i]),
    "ucy_ensemble_time_gemm": (_i, [_vp, _i, _pd, _pd]),
    "ucy_ensemble_time_phases": (_i, [_vp, _i, _pd]),
}

EXPORTED = tuple(_SIGS)
GEMM_TUNE_ENTRIES = 121   # UCY_GEMM_TUNE_ENTRIES of include/ucy.h
ENS_PHASES = ("start", "gemm 1", "post 1", "gemm 2", "post 2", "gemm 3", "post 3", "gemm 4",
              "post 4", "gemm 5", "post 5", "gemm 6", "finish", "control")  # UCY_ENS_PHASES


def load(path: str = LIB_PATH) -> C.CDLL:
    if not os.path.exists(path):
        raise ImportError(
            f"libucy.so not found at {path}: build it with "
            f"'make -C unicycle_amd/csrc' (there is no CPU fallback)")
    lib = C.CDLL(path)
    for name, (res, args) in _SIGS.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


_LIB = None


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        _LIB = load()
    return _LIB
