"""Python host over libfsim_hip, mirroring the reference's module interface.

The reference's main program (main.F90) drives three modules:
  * g_init / ptcl_init / lapl_init / circ_init   -> ``Discharge.__init__`` and
    ``ptcl_init()`` / ``init_gaussian()``
  * ptcl_step(g, ph, Em)                         -> ``ptcl_step()``
  * lapl_solve(g, ne1, ni1, nt1, sig, lerr)      -> ``lapl_solve()``
  * the accept / reject branches (main.F90:83-121) -> ``accept()`` / ``reject()``
  * one pass of the main loop (main.F90:68-121)  -> ``step()``
Module arrays (ni, ne, nt, nm levels 1..3 and ph) are read and written with
``get`` / ``set`` in the reference's (ny+2, nx+2) layout (i fastest).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib as _L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def descs(cfg: dict):
    """fs_grid_desc and fs_params of a configuration dictionary."""
    g = _L.GridDesc()
    for name, _ in _L.GridDesc._fields_:
        if name in cfg:
            setattr(g, name, cfg[name])
    p = _L.Params()
    for name, _ in _L.Params._fields_:
        if name in cfg:
            setattr(p, name, cfg[name])
    return g, p


class Discharge:
    SCALARS = ["t", "dt", "err_prev", "t_m", "Vd_pl", "Vd_mi", "Vd_or", "Ida", "Idc",
               "Vsrc", "ts"]

    def __init__(self, cfg: dict, handle=None, owner=None):
        """A discharge of configuration `cfg` (its own handle), or a view of an
        existing handle (an ensemble member: `handle`, kept alive by `owner`)."""
        L = _L.load()
        self.cfg = dict(cfg)
        self._g, self._p = descs(cfg)
        self._owner = owner
        if handle is None:
            h = C.c_void_p()
            _L.check(L.fs_create(C.byref(self._g), C.byref(self._p), C.byref(h)))
            self._owns = True
        else:
            h = C.c_void_p(handle)
            self._owns = False
        self.h = h
        self.nx, self.ny = int(cfg["nx"]), int(cfg["ny"])
        self.P = self.nx + 2

    def close(self):
        if getattr(self, "h", None) and getattr(self, "_owns", False):
            _L.load().fs_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- g_init outputs ----
    def grid(self) -> dict:
        nx, ny = self.nx, self.ny
        g = {k: np.zeros(n) for k, n in [("x", nx + 2), ("y", ny + 2), ("dx", nx + 1),
                                          ("dlx", nx), ("dy", ny + 1), ("dly", ny),
                                          ("r", ny + 2)]}
        tx = np.zeros(nx * ny, dtype=np.int32)
        ty = np.zeros(nx * ny, dtype=np.int32)
        ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
        _L.check(_L.load().fs_get_grid(self.h, _dp(g["x"]), _dp(g["y"]), _dp(g["dx"]),
                                       _dp(g["dlx"]), _dp(g["dy"]), _dp(g["dly"]),
                                       _dp(g["r"]), ip(tx), ip(ty)))
        g["type_x"] = tx.reshape(ny, nx)
        g["type_y"] = ty.reshape(ny, nx)
        return g

    # ---- module arrays ----
    def get(self, name: str, level: int = 1) -> np.ndarray:
        out = np.zeros((self.ny + 2, self.P))
        _L.check(_L.load().fs_get_field(self.h, _L.FIELDS[name], level, _dp(out), self.P))
        return out

    def set(self, name: str, level: int, arr) -> None:
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(self.ny + 2, self.P)
        _L.check(_L.load().fs_set_field(self.h, _L.FIELDS[name], level, _dp(a), self.P))

    def get_sig(self, which: str = "sig_pl") -> np.ndarray:
        """The dielectric wall's surface charge (rwall), x nodes 1..nx+2."""
        out = np.zeros(self.nx + 2)
        _L.check(_L.load().fs_get_field(self.h, _L.FS_SIG, _L.SIG_LEVELS[which], _dp(out), 0))
        return out

    def set_sig(self, which: str, arr) -> None:
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(self.nx + 2)
        _L.check(_L.load().fs_set_field(self.h, _L.FS_SIG, _L.SIG_LEVELS[which], _dp(a), 0))

    def circ_init(self, vset: float) -> None:
        """circ_init(Vset): source amplitude and circuit voltages (nondimensional)."""
        _L.check(_L.load().fs_circ_init(self.h, vset))

    def ptcl_init(self):
        _L.check(_L.load().fs_init_uniform(self.h))

    def init_gaussian(self, seed=0, peak=None, sigma=None, noise=0.01):
        from .configs import X0, X0_3
        peak = 1e14 * X0_3 if peak is None else peak
        sigma = 1e-3 / X0 if sigma is None else sigma
        _L.check(_L.load().fs_init_gaussian(self.h, int(seed), peak, sigma, noise))

    def scalars(self) -> dict:
        out = np.zeros(11)
        _L.check(_L.load().fs_get_scalars(self.h, _dp(out)))
        return dict(zip(self.SCALARS, out))

    def set_scalars(self, d: dict):
        cur = self.scalars()
        cur.update(d)
        a = np.array([cur[n] for n in self.SCALARS], dtype=np.float64)
        _L.check(_L.load().fs_set_scalars(self.h, _dp(a)))

    # ---- hot path ----
    def ptcl_step(self) -> _L.StepInfo:
        info = _L.StepInfo()
        _L.check(_L.load().fs_ptcl_step(self.h, C.byref(info)))
        return info

    def lapl_solve(self):
        lerr = C.c_int32(0)
        st = _L.SolveStats()
        _L.check(_L.load().fs_lapl_solve(self.h, C.byref(lerr), C.byref(st)))
        return lerr.value, st

    def accept(self):
        _L.check(_L.load().fs_accept(self.h))

    def reject(self):
        _L.check(_L.load().fs_reject(self.h))

    def step(self):
        """One pass of the main loop body (main.F90:74-121)."""
        if self.scalars()["ts"] > 1:
            self.ptcl_step()
        lerr, st = self.lapl_solve()
        if lerr == 0:
            self.accept()
        else:
            self.reject()
        return lerr, st

    def run(self, n_iters: int):
        _L.check(_L.load().fs_run(self.h, int(n_iters)))

    # ---- replay / log ----
    def set_replay(self, recs):
        r = np.ascontiguousarray(recs, dtype=np.float64).reshape(-1, _L.FS_REC_LEN)
        _L.check(_L.load().fs_set_replay(self.h, _dp(r), r.shape[0]))

    def stage_tile(self) -> int:
        """Tile height of the stage kernel in use (4, 8 or 16)."""
        return _L.check(_L.load().fs_stage_tile(self.h))

    def log(self) -> np.ndarray:
        L = _L.load()
        n = L.fs_log_count(self.h)
        out = np.zeros((max(n, 1), _L.FS_REC_LEN))
        L.fs_get_log(self.h, _dp(out), n)
        return out[:n]

    def clear_log(self):
        _L.check(_L.load().fs_clear_log(self.h))

    # ---- sub-steps for kernel parity ----
    def rhs_sweep(self, stage: int):
        _L.check(_L.load().fs_rhs_sweep(self.h, stage))

    def circ_step(self, stage: int):
        _L.check(_L.load().fs_circ_step(self.h, stage))

    def lapl_assemble(self):
        N = self.nx * self.ny
        b = np.zeros(N)
        A5 = np.zeros((N, 5))
        _L.check(_L.load().fs_lapl_assemble(self.h, _dp(b), _dp(A5)))
        return b.reshape(self.ny, self.nx), A5.reshape(self.ny, self.nx, 5)

    def set_solver(self, solver: int):
        """Poisson preconditioner of later solves: 0 Jacobi, 1 multigrid."""
        _L.check(_L.load().fs_set_solver(self.h, int(solver)))

    def set_timing(self, on: bool):
        _L.check(_L.load().fs_set_timing(self.h, int(on)))

    def time_kernel(self, which: int, n: int = 50) -> float:
        out = np.zeros(1)
        _L.check(_L.load().fs_time_kernel(self.h, int(which), int(n), _dp(out)))
        return float(out[0])

    def timers(self) -> np.ndarray:
        out = np.zeros(4)
        _L.check(_L.load().fs_get_timers(self.h, _dp(out)))
        return out


def copy_bandwidth(nbytes: int = 1 << 30, reps: int = 20) -> float:
    """Measured device copy bandwidth (GB/s, read + write bytes)."""
    out = np.zeros(1)
    _L.check(_L.load().fs_copy_bench(int(nbytes), int(reps), _dp(out)))
    return float(out[0])


def constants(p_torr=10.0, gam=0.1) -> np.ndarray:
    out = np.zeros(40)
    _L.check(_L.load().fs_constants(p_torr, gam, _dp(out)))
    return out


def eval_fits(T, p_torr=10.0) -> np.ndarray:
    T = np.ascontiguousarray(T, dtype=np.float64)
    out = np.zeros((T.size, 12))
    _L.check(_L.load().fs_eval_fits(p_torr, T.size, _dp(T), _dp(out)))
    return out


def eval_flux(args) -> np.ndarray:
    a = np.ascontiguousarray(args, dtype=np.float64).reshape(-1, 7)
    out = np.zeros(a.shape[0])
    _L.check(_L.load().fs_eval_flux(a.shape[0], _dp(a), _dp(out)))
    return out


def eval_dfdph(args) -> np.ndarray:
    a = np.ascontiguousarray(args, dtype=np.float64).reshape(-1, 7)
    out = np.zeros(a.shape[0])
    _L.check(_L.load().fs_eval_dfdph(a.shape[0], _dp(a), _dp(out)))
    return out
