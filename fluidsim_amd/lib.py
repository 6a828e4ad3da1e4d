"""ctypes binding of libfsim_hip.so (declarations in include/fsim.h).

There is no fallback: if the HIP library is missing or fails to load, importing
this module raises, so nothing can silently run on the CPU.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfsim_hip.so")

FS_NI, FS_NE, FS_NT, FS_NM, FS_PH, FS_KI, FS_KE, FS_KT, FS_KM, FS_SIG = range(10)
FIELDS = {"ni": FS_NI, "ne": FS_NE, "nt": FS_NT, "nm": FS_NM, "ph": FS_PH,
          "ki": FS_KI, "ke": FS_KE, "kt": FS_KT, "km": FS_KM}
SIG_LEVELS = {"sig_pl": 1, "sig_mi": 2, "sig_or": 3}
FS_REC_LEN = 10
STATUS = {0: "ok", 1: "not converged", -1: "bad argument", -2: "negative density",
          -3: "NaN in stage slope", -4: "minimum timestep", -5: "NaN in dFdPh",
          -6: "HIP error", -7: "replay exhausted"}

EXPORTS = [
    "fs_create", "fs_destroy", "fs_get_grid", "fs_set_field", "fs_get_field",
    "fs_init_uniform", "fs_init_gaussian", "fs_circ_init", "fs_stage_tile", "fs_get_scalars",
    "fs_set_scalars",
    "fs_ptcl_step", "fs_lapl_solve", "fs_accept", "fs_reject", "fs_run", "fs_set_replay",
    "fs_log_count", "fs_get_log", "fs_clear_log", "fs_rhs_sweep", "fs_circ_step",
    "fs_lapl_assemble", "fs_constants", "fs_eval_fits", "fs_eval_flux", "fs_eval_dfdph",
    "fs_eval_flux_noalloc",
    "fs_get_timers", "fs_set_timing", "fs_set_solver", "fs_time_kernel", "fs_copy_bench",
    "fs_ensemble_create", "fs_ensemble_destroy", "fs_ensemble_member", "fs_ensemble_run",
    "fs_ensemble_stage_tile",
    "fs_learn_create", "fs_learn_destroy", "fs_learn_nparams", "fs_learn_get_weights",
    "fs_learn_set_weights", "fs_learn_samples", "fs_learn_harvest", "fs_learn_train",
    "fs_learn_loss_grad", "fs_learn_attach", "fs_last_error",
]


class GridDesc(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("grid_x", C.c_int32),
                ("grid_y", C.c_int32), ("ratio_x", C.c_double), ("ratio_y", C.c_double),
                ("l", C.c_double), ("w", C.c_double), ("ew", C.c_double), ("gap", C.c_double),
                ("cyl", C.c_int32), ("rwall", C.c_int32)]


class Params(C.Structure):
    _fields_ = [("p_torr", C.c_double), ("gam", C.c_double), ("vset", C.c_double),
                ("res_ohm", C.c_double), ("rf", C.c_int32), ("circuit", C.c_int32),
                ("t_rise", C.c_double), ("t_width", C.c_double), ("t_fall", C.c_double),
                ("dt0", C.c_double), ("lapl_rtol", C.c_double), ("lapl_maxit", C.c_int32),
                ("lapl_chunk", C.c_int32), ("debug_keep_k", C.c_int32), ("lapl_solver", C.c_int32),
                ("use_graphs", C.c_int32), ("lapl_warm", C.c_int32), ("stage_tile", C.c_int32)]


class StepInfo(C.Structure):
    _fields_ = [("err_cur", C.c_double), ("dt_stage", C.c_double), ("dt_next", C.c_double),
                ("t_m", C.c_double), ("Vd_pl", C.c_double), ("Ida", C.c_double),
                ("Idc", C.c_double), ("Vsrc", C.c_double), ("attempts", C.c_int32),
                ("status", C.c_int32)]


class SolveStats(C.Structure):
    _fields_ = [("iters", C.c_int32), ("lerr", C.c_int32), ("resid", C.c_double),
                ("bnorm", C.c_double), ("seconds", C.c_double)]


class FsError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"libfsim_hip status {code} ({STATUS.get(code, '?')}): {msg}")
        self.code = code


_lib = None


def load():
    """Load libfsim_hip.so (raises if it is absent: the product has no CPU path)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing; build it with __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    P = C.c_void_p
    D = C.POINTER(C.c_double)
    I = C.POINTER(C.c_int32)
    F = C.POINTER(C.c_float)
    sig = {
        "fs_create": [C.POINTER(GridDesc), C.POINTER(Params), C.POINTER(P)],
        "fs_destroy": [P],
        "fs_get_grid": [P, D, D, D, D, D, D, D, I, I],
        "fs_set_field": [P, C.c_int32, C.c_int32, D, C.c_int32],
        "fs_get_field": [P, C.c_int32, C.c_int32, D, C.c_int32],
        "fs_init_uniform": [P],
        "fs_init_gaussian": [P, C.c_uint64, C.c_double, C.c_double, C.c_double],
        "fs_circ_init": [P, C.c_double],
        "fs_stage_tile": [P],
        "fs_get_scalars": [P, D],
        "fs_set_scalars": [P, D],
        "fs_ptcl_step": [P, C.POINTER(StepInfo)],
        "fs_lapl_solve": [P, I, C.POINTER(SolveStats)],
        "fs_accept": [P],
        "fs_reject": [P],
        "fs_run": [P, C.c_int32],
        "fs_set_replay": [P, D, C.c_int32],
        "fs_log_count": [P],
        "fs_get_log": [P, D, C.c_int32],
        "fs_clear_log": [P],
        "fs_rhs_sweep": [P, C.c_int32],
        "fs_circ_step": [P, C.c_int32],
        "fs_lapl_assemble": [P, D, D],
        "fs_constants": [C.c_double, C.c_double, D],
        "fs_eval_fits": [C.c_double, C.c_int32, D, D],
        "fs_eval_flux": [C.c_int32, D, D],
        "fs_eval_dfdph": [C.c_int32, D, D],
        "fs_eval_flux_noalloc": [C.c_int32, D, D, C.c_int32],
        "fs_get_timers": [P, D],
        "fs_set_timing": [P, C.c_int32],
        "fs_set_solver": [P, C.c_int32],
        "fs_time_kernel": [P, C.c_int32, C.c_int32, D],
        "fs_copy_bench": [C.c_int64, C.c_int32, D],
        "fs_ensemble_create": [C.c_int32, C.POINTER(GridDesc), C.POINTER(Params), C.POINTER(P)],
        "fs_ensemble_destroy": [P],
        "fs_ensemble_member": [P, C.c_int32],
        "fs_ensemble_run": [P, C.c_int32],
        "fs_ensemble_stage_tile": [P],
        "fs_learn_create": [C.c_int32, C.c_int32, C.c_int32, C.c_uint64, C.POINTER(P)],
        "fs_learn_destroy": [P],
        "fs_learn_nparams": [P],
        "fs_learn_get_weights": [P, F],
        "fs_learn_set_weights": [P, F],
        "fs_learn_samples": [P],
        "fs_learn_harvest": [P, P, I],
        "fs_learn_train": [P, C.c_int32, C.c_int32, C.c_float, C.c_uint64, D],
        "fs_learn_loss_grad": [P, C.c_int32, C.c_int32, C.c_int32, F, F, F, D, F, F],
        "fs_learn_attach": [P, P],
        "fs_last_error": [],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    L.fs_last_error.restype = C.c_char_p
    L.fs_learn_nparams.restype = C.c_int64
    L.fs_ensemble_member.restype = C.c_void_p
    _lib = L
    return L


def check(code: int) -> int:
    if code < 0:
        raise FsError(code, load().fs_last_error().decode())
    return code
