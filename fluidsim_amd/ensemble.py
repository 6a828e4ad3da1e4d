"""Ensembles of independent discharges (C-ABI fs_ensemble_*, include/fsim.h).

Every member is an ordinary discharge -- own grid, physics, adaptive dt,
controller and Poisson convergence -- and the ensemble advances all of them
with launches that each cover every member.  A member's fields after
``run(k)`` are bitwise those of the same discharge run alone for k passes.
"""
from __future__ import annotations

import ctypes as C

from . import lib as _L
from .discharge import Discharge, descs


class Ensemble:
    def __init__(self, cfgs):
        L = _L.load()
        n = len(cfgs)
        gs = (_L.GridDesc * n)()
        ps = (_L.Params * n)()
        for i, c in enumerate(cfgs):
            gs[i], ps[i] = descs(c)
        h = C.c_void_p()
        _L.check(L.fs_ensemble_create(n, gs, ps, C.byref(h)))
        self.h = h
        self.cfgs = [dict(c) for c in cfgs]
        self.members = []
        for i, c in enumerate(cfgs):
            mh = L.fs_ensemble_member(h, i)
            if not mh:
                raise _L.FsError(-1, L.fs_last_error().decode())
            self.members.append(Discharge(c, handle=mh, owner=self))

    def __len__(self):
        return len(self.members)

    def __getitem__(self, i) -> Discharge:
        return self.members[i]

    def stage_tile(self) -> int:
        """Tile height of the batched stage launches (4, 8 or 16)."""
        return _L.check(_L.load().fs_ensemble_stage_tile(self.h))

    def run(self, passes: int) -> None:
        _L.check(_L.load().fs_ensemble_run(self.h, int(passes)))

    def close(self):
        if self.h:
            for m in self.members:
                m.h = None
            _L.load().fs_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
