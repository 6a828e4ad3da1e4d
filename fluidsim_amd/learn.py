"""The learned Poisson preconditioner (C-ABI fs_learn_*, declared in
include/fsim.h).

A 4-level U-Net (16/32/48/64 channels, 3x3 kernels, BF16 weights and
activations, FP32 accumulation on the MFMA units) is the coarse solver of a
two-grid preconditioner inside flexible CG (``lapl_solver = 2``).  It is
trained on the device, by the library's own forward/backward kernels, with the
A-norm energy loss over coarse residuals harvested from the simulator's
multigrid-preconditioned solves.  The plain preconditioners stay available
(``Discharge.set_solver(0 | 1)``) and every solve, learned or not, runs to the
same FP64 residual.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib as _L


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class LearnedPreconditioner:
    def __init__(self, capacity: int, H: int, W: int, seed: int = 0):
        """Training set of `capacity` samples on an H x W coarse grid
        (multiples of 8; the coarse grid of an n x m discharge is n/2 x m/2)."""
        L = _L.load()
        h = C.c_void_p()
        _L.check(L.fs_learn_create(int(capacity), int(H), int(W), int(seed), C.byref(h)))
        self.h, self.H, self.W = h, H, W

    def close(self):
        if self.h:
            _L.load().fs_learn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def nparams(self) -> int:
        return int(_L.load().fs_learn_nparams(self.h))

    @property
    def samples(self) -> int:
        return int(_L.load().fs_learn_samples(self.h))

    def weights(self) -> np.ndarray:
        w = np.zeros(self.nparams, dtype=np.float32)
        _L.check(_L.load().fs_learn_get_weights(self.h, _fp(w)))
        return w

    def set_weights(self, w) -> None:
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.size == self.nparams
        _L.check(_L.load().fs_learn_set_weights(self.h, _fp(w)))

    def harvest(self, discharge) -> int:
        """Samples from one multigrid solve of the discharge's current system."""
        n = C.c_int32(0)
        _L.check(_L.load().fs_learn_harvest(self.h, discharge.h, C.byref(n)))
        return n.value

    def train(self, steps: int, batch: int = 8, lr: float = 1e-3, seed: int = 0) -> np.ndarray:
        loss = np.zeros(max(steps, 1))
        _L.check(_L.load().fs_learn_train(self.h, int(steps), int(batch), float(lr), int(seed),
                                         _dp(loss)))
        return loss[:steps]

    def loss_grad(self, in5, rhat, op):
        """Loss, parameter gradient and e_hat of a host batch:
        in5 [B, H, W, 5], rhat [B, H, W], op [B, 3, H, W]."""
        in5 = np.ascontiguousarray(in5, dtype=np.float32)
        rhat = np.ascontiguousarray(rhat, dtype=np.float32)
        op = np.ascontiguousarray(op, dtype=np.float32)
        B, H, W, _ = in5.shape
        loss = np.zeros(1)
        grad = np.zeros(self.nparams, dtype=np.float32)
        out = np.zeros((B, H, W), dtype=np.float32)
        _L.check(_L.load().fs_learn_loss_grad(self.h, B, H, W, _fp(in5), _fp(rhat), _fp(op),
                                             _dp(loss), _fp(grad), _fp(out)))
        return float(loss[0]), grad, out

    @staticmethod
    def detach(discharge) -> None:
        """Learned solves of the discharge fall back to a zero correction on the
        learned level (the method without the network)."""
        _L.check(_L.load().fs_learn_attach(discharge.h, None))
        discharge._net = None

    def attach(self, discharge) -> None:
        """Use this network in the discharge's solves with lapl_solver = 2."""
        _L.check(_L.load().fs_learn_attach(discharge.h, self.h))
        discharge._net = self  # keep the network alive while attached


def harvest_runs(net: LearnedPreconditioner, n: int, seeds, samples: int, warm: int = 5,
                 every: int = 2, config=None) -> float:
    """Fill the training set with `samples` residuals from seeded runs of an
    n x n discharge (the 512^2 pulse workload by default): after `warm` passes,
    one multigrid solve is harvested every `every` passes.  Returns seconds."""
    import time

    from . import configs
    from .discharge import Discharge

    config = config or configs.config1
    t0 = time.perf_counter()
    target0 = net.samples
    for k, seed in enumerate(seeds):
        d = Discharge(config(n))
        d.init_gaussian(seed=int(seed))
        d.run(warm)
        goal = target0 + samples * (k + 1) // len(seeds)
        while net.samples < goal:
            d.run(every)
            if net.harvest(d) == 0:
                break
    return time.perf_counter() - t0


def harvest_ensemble(net: LearnedPreconditioner, members: int, seeds, samples: int,
                     warm: int = 4, every: int = 2) -> dict:
    """Add `samples` residuals from the multigrid solves of the parameter-sweep
    ensemble (sides 64..256, DC 150..400 V, 0.5..10 Torr): for each seed an
    ensemble of `members` discharges runs `warm` passes, then one solve of
    every member is harvested every `every` passes.  Members whose learned
    level does not fit the network's grid (a side above 2 x its width) are
    skipped.  Returns seconds, members used and skipped."""
    import time

    from . import configs
    from .ensemble import Ensemble

    t0 = time.perf_counter()
    target0 = net.samples
    used, skipped = set(), set()
    for k, seed in enumerate(seeds):
        cfgs = configs.config3(members, base_seed=int(seed) * members)
        E = Ensemble(cfgs)
        for m, c in zip(E.members, cfgs):
            m.init_gaussian(seed=c["seed"])
        E.run(warm)
        goal = target0 + samples * (k + 1) // len(seeds)
        while net.samples < goal:
            E.run(every)
            got = 0
            for i, m in enumerate(E.members):
                if (k, i) in skipped:
                    continue
                try:
                    got += net.harvest(m)
                    used.add((k, i))
                except _L.FsError:
                    skipped.add((k, i))
                if net.samples >= goal:
                    break
            if got == 0:
                break
        E.close()
    return {"seconds": time.perf_counter() - t0, "members_used": len(used),
            "members_skipped": len(skipped)}


def train_schedule(net: LearnedPreconditioner, phases=((1500, 16, 1e-3), (1500, 16, 2.5e-4)),
                   seed: int = 0):
    """Adam phases of (steps, batch, learning rate); returns (losses, seconds)."""
    import time

    t0 = time.perf_counter()
    losses = [net.train(s, batch=b, lr=lr, seed=seed + k) for k, (s, b, lr) in enumerate(phases)]
    return np.concatenate(losses), time.perf_counter() - t0


def kernels_workload(n: int = 512, seeds: int = 8, steps: int = 200, reps: int = 20) -> dict:
    """A short workload of the learned preconditioner's own kernels that does
    not advance any discharge: residuals harvested from the multigrid solves
    of Gaussian initial states of `seeds` seeds, `steps` Adam steps of the
    device training (forward, backward, weight gradients, Adam), then `reps`
    network applications and learned flexible-CG iterations timed with HIP
    events.  Returns the timings (ms) and the training summary."""
    import time

    from . import configs
    from .discharge import Discharge

    d = Discharge(configs.config1(n))
    net = LearnedPreconditioner(max(64, 8 * seeds), 64, 64, seed=1)
    t0 = time.perf_counter()
    for s in range(seeds):
        d.init_gaussian(seed=s)
        net.harvest(d)
    t1 = time.perf_counter()
    losses = net.train(steps, batch=8, lr=1e-3, seed=2)
    t2 = time.perf_counter()
    net.attach(d)
    d.set_solver(2)
    out = {"n": n, "samples": net.samples, "harvest_s": t1 - t0, "train_s": t2 - t1,
           "train_steps": steps, "loss_first": float(losses[0]), "loss_last": float(losses[-1]),
           "network_apply_ms": d.time_kernel(15, reps), "learned_iteration_ms": d.time_kernel(16, reps)}
    d.set_solver(1)
    d.close()
    net.close()
    return out


if __name__ == "__main__":
    import argparse
    import json

    ap = argparse.ArgumentParser(description=kernels_workload.__doc__)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--seeds", type=int, default=8)
    ap.add_argument("--steps", type=int, default=200)
    a = ap.parse_args()
    print(json.dumps(kernels_workload(a.n, a.seeds, a.steps)))
