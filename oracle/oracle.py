"""ctypes front end of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
bench.py's cpu_baseline leg, never by the product package.

The oracle restates the reference Fortran hot path in C (fsim_oracle.c, with
per-routine citations of ptcl_props.F90, ptcl_lib.F90, circ_lib.F90,
lapl_lib.F90, props.F90 and main.F90).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "_build", "liboracle.so")

FIELDS = {"ni": 0, "ne": 1, "nt": 2, "nm": 3, "ph": 4, "ki": 5, "ke": 6, "kt": 7, "km": 8}
OREC_LEN = 10
NCONST = 40
CONST_NAMES = [
    "pi", "eps0", "mu0", "e", "kb", "me", "x0", "ph0", "t0", "wp02", "w2", "Tg", "p",
    "ninf", "n_zero", "n_init", "mi", "mui", "vi", "Di", "Dm", "k_r", "k_mp", "k_2q",
    "k_3q", "H_ir", "H_ex", "H_si", "H_sc", "gam", "log5119", "log5667", "log75", "ftol",
    "Cap", "x0_3", "c0", "kTe", "nmin_t", "x0_6",
]


class OCfg(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("grid_x", C.c_int32), ("grid_y", C.c_int32),
        ("ratio_x", C.c_double), ("ratio_y", C.c_double),
        ("l", C.c_double), ("w", C.c_double), ("ew", C.c_double), ("gap", C.c_double),
        ("cyl", C.c_int32), ("rf", C.c_int32),
        ("p_torr", C.c_double), ("gam", C.c_double), ("vset", C.c_double),
        ("res_ohm", C.c_double), ("t_rise", C.c_double), ("t_width", C.c_double),
        ("t_fall", C.c_double), ("circuit", C.c_int32), ("solver", C.c_int32),
        ("dt0", C.c_double), ("rtol", C.c_double), ("maxit", C.c_int32),
        ("nthreads", C.c_int32), ("rwall", C.c_int32),
    ]


def build(force: bool = False) -> str:
    if force or not os.path.exists(_SO):
        subprocess.check_call(["make", "-s", "-C", _HERE])
    return _SO


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = C.CDLL(_SO)
        P = C.c_void_p
        D = C.POINTER(C.c_double)
        I = C.POINTER(C.c_int32)
        L.o_create.restype = P
        L.o_create.argtypes = [C.POINTER(OCfg)]
        L.o_destroy.argtypes = [P]
        L.o_dims.argtypes = [P, I]
        L.o_get_grid.argtypes = [P, D, D, D, D, D, D, D, I, I]
        L.o_get_field.argtypes = [P, C.c_int, C.c_int, D]
        L.o_set_field.argtypes = [P, C.c_int, C.c_int, D]
        L.o_init_uniform.argtypes = [P]
        L.o_init_gaussian.argtypes = [P, C.c_uint64, C.c_double, C.c_double, C.c_double]
        L.o_get_scalars.argtypes = [P, D]
        L.o_set_scalars.argtypes = [P, D]
        L.o_rhs_sweep.argtypes = [P, C.c_int]
        L.o_circ_step.argtypes = [P, C.c_int]
        L.o_sfc_step.argtypes = [P, C.c_int]
        L.o_get_sig.argtypes = [P, C.c_int, D]
        L.o_set_sig.argtypes = [P, C.c_int, D]
        L.o_rk_all.argtypes = [P, C.c_int, D]
        L.o_ptcl_step.argtypes = [P]
        L.o_lapl_assemble.argtypes = [P, D, D, D, D, D]
        L.o_lapl_solve.argtypes = [P, C.POINTER(C.c_int), D]
        L.o_accept.argtypes = [P]
        L.o_reject.argtypes = [P]
        L.o_run.argtypes = [P, C.c_int, D]
        L.o_set_replay.argtypes = [P, D, C.c_int]
        L.o_log_count.argtypes = [P]
        L.o_get_log.argtypes = [P, D, C.c_int]
        L.o_clear_log.argtypes = [P]
        L.o_consts.argtypes = [C.c_double, C.c_double, D]
        L.o_eval_fits.argtypes = [C.c_double, C.c_int, D, D]
        L.o_eval_flux.argtypes = [C.c_int, D, D]
        L.o_eval_dfdph.argtypes = [C.c_int, D, D]
        L.o_gaussian_field.argtypes = [C.POINTER(OCfg), C.c_uint64, C.c_double, C.c_double,
                                       C.c_double, C.c_int, D]
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def make_cfg(d: dict) -> OCfg:
    c = OCfg()
    for name, _ in OCfg._fields_:
        if name in d:
            setattr(c, name, d[name])
    return c


def consts(p_torr=10.0, gam=0.1) -> dict:
    out = np.zeros(NCONST)
    lib().o_consts(p_torr, gam, _dp(out))
    return dict(zip(CONST_NAMES, out))


def consts_array(p_torr=10.0, gam=0.1) -> np.ndarray:
    out = np.zeros(NCONST)
    lib().o_consts(p_torr, gam, _dp(out))
    return out


def eval_fits(T, p_torr=10.0) -> np.ndarray:
    T = np.ascontiguousarray(T, dtype=np.float64)
    out = np.zeros((T.size, 12))
    lib().o_eval_fits(p_torr, T.size, _dp(T), _dp(out))
    return out


def eval_flux(args) -> np.ndarray:
    a = np.ascontiguousarray(args, dtype=np.float64).reshape(-1, 7)
    out = np.zeros(a.shape[0])
    lib().o_eval_flux(a.shape[0], _dp(a), _dp(out))
    return out


def eval_dfdph(args) -> np.ndarray:
    a = np.ascontiguousarray(args, dtype=np.float64).reshape(-1, 7)
    out = np.zeros(a.shape[0])
    lib().o_eval_dfdph(a.shape[0], _dp(a), _dp(out))
    return out


class Oracle:
    """One simulation instance of the CPU restatement (reference layout)."""

    def __init__(self, cfg: dict):
        self.cfg = dict(cfg)
        self._c = make_cfg(cfg)
        self.h = lib().o_create(C.byref(self._c))
        if not self.h:
            raise ValueError("oracle: invalid configuration")
        d = np.zeros(4, dtype=np.int32)
        lib().o_dims(self.h, _ip(d))
        self.bx, self.by, self.P = int(d[0]), int(d[1]), int(d[2])

    def __del__(self):
        if getattr(self, "h", None):
            lib().o_destroy(self.h)
            self.h = None

    # -- grid --
    def grid(self) -> dict:
        bx, by = self.bx, self.by
        g = {k: np.zeros(n) for k, n in [("x", bx + 2), ("y", by + 2), ("dx", bx + 1),
                                          ("dlx", bx), ("dy", by + 1), ("dly", by),
                                          ("r", by + 2)]}
        tx = np.zeros(bx * by, dtype=np.int32)
        ty = np.zeros(bx * by, dtype=np.int32)
        lib().o_get_grid(self.h, _dp(g["x"]), _dp(g["y"]), _dp(g["dx"]), _dp(g["dlx"]),
                         _dp(g["dy"]), _dp(g["dly"]), _dp(g["r"]), _ip(tx), _ip(ty))
        g["type_x"] = tx.reshape(by, bx)
        g["type_y"] = ty.reshape(by, bx)
        return g

    # -- fields: arrays shaped (by+2, bx+2), i fastest (reference column-major) --
    def get(self, name: str, level: int = 1) -> np.ndarray:
        out = np.zeros((self.by + 2, self.P))
        if lib().o_get_field(self.h, FIELDS[name], level, _dp(out)) != 0:
            raise ValueError(name)
        return out

    def set(self, name: str, level: int, arr) -> None:
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(self.by + 2, self.P)
        if lib().o_set_field(self.h, FIELDS[name], level, _dp(a)) != 0:
            raise ValueError(name)

    def init_uniform(self):
        lib().o_init_uniform(self.h)

    def init_gaussian(self, seed=0, peak=None, sigma=None, noise=0.01):
        k = consts(self.cfg.get("p_torr", 10.0))
        if peak is None:
            peak = 1e14 * k["x0_3"]  # 1e8 cm^-3
        if sigma is None:
            sigma = 1e-3 / k["x0"]  # 0.1 cm
        lib().o_init_gaussian(self.h, seed, peak, sigma, noise)

    def scalars(self) -> dict:
        out = np.zeros(11)
        lib().o_get_scalars(self.h, _dp(out))
        names = ["t", "dt", "err_prev", "t_m", "Vd_pl", "Vd_mi", "Vd_or", "Ida", "Idc",
                 "Vsrc", "ts"]
        return dict(zip(names, out))

    def set_scalars(self, d: dict):
        cur = self.scalars()
        cur.update(d)
        names = ["t", "dt", "err_prev", "t_m", "Vd_pl", "Vd_mi", "Vd_or", "Ida", "Idc",
                 "Vsrc", "ts"]
        a = np.array([cur[n] for n in names], dtype=np.float64)
        lib().o_set_scalars(self.h, _dp(a))

    # -- hot path --
    def rhs_sweep(self, stage: int):
        lib().o_rhs_sweep(self.h, stage)

    def circ_step(self, stage: int):
        lib().o_circ_step(self.h, stage)

    def sfc_step(self, stage: int):
        lib().o_sfc_step(self.h, stage)

    SIG = {"sig_pl": 0, "sig_mi": 1, "sig_or": 2}

    def get_sig(self, which: str = "sig_pl") -> np.ndarray:
        """Wall surface charge at x-nodes 1..bx+2 (ghosts included)."""
        out = np.zeros(self.bx + 2)
        lib().o_get_sig(self.h, self.SIG[which], _dp(out))
        return out

    def set_sig(self, which: str, arr) -> None:
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(self.bx + 2)
        lib().o_set_sig(self.h, self.SIG[which], _dp(a))

    def rk_all(self, stage: int) -> np.ndarray:
        e = np.zeros(4)
        lib().o_rk_all(self.h, stage, _dp(e))
        return e

    def ptcl_step(self) -> int:
        return lib().o_ptcl_step(self.h)

    def lapl_assemble(self):
        ne, ni, nt = (np.ascontiguousarray(self.get(n, 1)) for n in ("ne", "ni", "nt"))
        N = self.bx * self.by
        b = np.zeros(N)
        A5 = np.zeros((N, 5))
        st = lib().o_lapl_assemble(self.h, _dp(ne), _dp(ni), _dp(nt), _dp(b), _dp(A5))
        return st, b.reshape(self.by, self.bx), A5.reshape(self.by, self.bx, 5)

    def lapl_solve(self):
        lerr = C.c_int(0)
        stats = np.zeros(3)
        st = lib().o_lapl_solve(self.h, C.byref(lerr), _dp(stats))
        return st, lerr.value, stats

    def accept(self):
        lib().o_accept(self.h)

    def reject(self):
        lib().o_reject(self.h)

    def run(self, n_iters: int):
        t = np.zeros(3)
        st = lib().o_run(self.h, n_iters, _dp(t))
        if st < 0:
            raise RuntimeError(f"oracle fatal status {st}")
        return t

    def set_replay(self, recs: np.ndarray):
        r = np.ascontiguousarray(recs, dtype=np.float64).reshape(-1, OREC_LEN)
        lib().o_set_replay(self.h, _dp(r), r.shape[0])

    def log(self) -> np.ndarray:
        n = lib().o_log_count(self.h)
        out = np.zeros((max(n, 1), OREC_LEN))
        lib().o_get_log(self.h, _dp(out), n)
        return out[:n]

    def clear_log(self):
        lib().o_clear_log(self.h)
