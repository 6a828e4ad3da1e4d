"""GPU parity: libfsim_hip (through its C ABI) against the CPU oracle.

Tolerances: the only intended differences are the device exp/log/pow versus
the host libm (<= 1-2 ulp each), amplified where the reference itself is
ill-conditioned (SG cancellation 1 - exp(-arg) for small arg), reduction order
(error sums, Poisson dot products) and the Poisson solution, which both sides
compute to near its rounding floor.
"""
import numpy as np
import pytest

from conftest import oracle_cfg, rel_l2

pytestmark = pytest.mark.gpu


def _pair(O, cfg, seed=1, solver=0):
    from fluidsim_amd import Discharge

    d = Discharge(cfg)
    o = O.Oracle(oracle_cfg(cfg, solver=solver))
    d.init_gaussian(seed=seed)
    o.init_gaussian(seed=seed)
    return d, o


def test_fits_match_oracle(O):
    from fluidsim_amd import eval_fits

    k = O.consts(10.0)
    v = np.concatenate([np.geomspace(0.05, 400.0, 3001),
                        [0.234, 0.9611, 1.37, 0.5, 16.0, 157.0, 5.119, 5.667, 7.5]])
    T = v / k["ph0"]
    for p in (0.5, 2.0, 10.0):
        a = eval_fits(T, p)
        b = O.eval_fits(T, p)
        # a 1-ulp log difference moves s = poly(x) by |ds/dx|*x*eps (up to ~1e4*eps
        # for the 11th-degree De/Dt fits), hence exp(s) by ~1e-12 relative
        for col in range(10):
            r = np.abs(a[:, col] - b[:, col]) / np.abs(b[:, col])
            assert r.max() < 1e-11, (p, col, r.max())
        # the stage kernel's own mobility and diffusion (folded conversions,
        # its exp and log) against the reference's get_mue and get_De
        for col, ref in ((10, 0), (11, 2)):
            r = np.abs(a[:, col] - b[:, ref]) / np.abs(b[:, ref])
            assert r.max() < 1e-11, (p, col, r.max())


def test_sg_flux_and_dfdph_match_oracle(O):
    from fluidsim_amd import eval_dfdph, eval_flux

    rng = np.random.default_rng(7)
    n = 20000
    E = rng.standard_normal(n) * 10.0 ** rng.uniform(-14, 5, n)
    h = rng.uniform(0.2, 3.0, n)
    q = rng.choice([-1.0, 1.0], n)
    mu = 10.0 ** rng.uniform(-3, 3, n)
    D = 10.0 ** rng.uniform(-2, 4, n)
    nl = 10.0 ** rng.uniform(-4, 8, n)
    nr = 10.0 ** rng.uniform(-4, 8, n)
    args = np.stack([E, h, q, mu, D, nl, nr], 1)
    arg = np.abs(q * mu * E * h / D)
    # SG cancellation: |1 - exp(-arg)| carries the device/host exp ulp as eps/arg;
    # dFdPh adds two such terms of size n/arg that cancel to O(n): eps/arg^2
    am = np.minimum(arg, 1.0)
    tols = (1e-13 + 8e-16 / am, 1e-13 + 8e-16 / am ** 2)
    for (fn_d, fn_o), tol in zip(((eval_flux, O.eval_flux), (eval_dfdph, O.eval_dfdph)), tols):
        a, b = fn_d(args), fn_o(args)
        # where 1 - exp(-a) rounds to 0 both sides give inf/NaN (the reference
        # stops there: lapl_lib.F90:632-635); the same inputs must be non-finite
        fin = np.isfinite(b)
        assert np.array_equal(np.isfinite(a), fin)
        scale = np.abs(b) + np.abs(q * mu * E) * (nl + nr) + D * (nl + nr) / h
        err = np.abs(a - b)[fin] / scale[fin]
        assert np.all(err <= tol[fin]), float((err / tol[fin]).max())


def _warm(O, cfg, steps=4, seed=1):
    """Oracle run to a nontrivial state, then the same state in both."""
    d, o = _pair(O, cfg, seed=seed)
    o.run(steps)
    for name in ("ni", "ne", "nt", "nm"):
        for lev in (1, 2, 3):
            d.set(name, lev, o.get(name, lev))
    d.set("ph", 1, o.get("ph"))
    s = o.scalars()
    d.set_scalars({k: s[k] for k in d.SCALARS})
    return d, o


@pytest.mark.parametrize("shape", [(64, 48), (17, 9)])
@pytest.mark.parametrize("stage", [1, 3, 5])
def test_rhs_sweep_matches_oracle(O, shape, stage):
    from fluidsim_amd import configs

    cfg = configs.base(*shape, debug=1, rtol=1e-13)
    d, o = _warm(O, cfg)
    d.rhs_sweep(stage)
    o.rhs_sweep(stage)
    for name in ("ki", "ke", "kt", "km"):
        a = d.get(name, stage)[1:-1, 1:-1]
        b = o.get(name, stage)[1:-1, 1:-1]
        assert rel_l2(a, b) < 1e-11, (name, rel_l2(a, b))


def test_lapl_assembly_matches_oracle(O):
    from fluidsim_amd import configs

    cfg = configs.base(64, 48, debug=1, rtol=1e-13)
    d, o = _warm(O, cfg)
    b_d, A_d = d.lapl_assemble()
    st, b_o, A_o = o.lapl_assemble()
    assert st == 0
    assert rel_l2(b_d, b_o) < 1e-10, rel_l2(b_d, b_o)
    for c in range(5):
        assert rel_l2(A_d[..., c], A_o[..., c]) < 1e-12, (c, rel_l2(A_d[..., c], A_o[..., c]))


def _compare(d, o, level=1):
    out = {}
    for name in ("ni", "ne", "nt", "nm"):
        out[name] = rel_l2(d.get(name, level)[1:-1, 1:-1], o.get(name, level)[1:-1, 1:-1])
    out["ph"] = rel_l2(d.get("ph")[1:-1, 1:-1], o.get("ph")[1:-1, 1:-1])
    return out


def test_one_step_parity(O):
    """From an identical state, one main-loop pass (5 stages, circuit,
    controller, Poisson, accept) with the oracle's dt decisions replayed."""
    from fluidsim_amd import configs

    cfg = configs.config0(rtol=1e-13)
    d, o = _warm(O, cfg, steps=6)
    o.clear_log()
    o.run(1)
    d.set_replay(o.log())
    d.run(1)
    errs = _compare(d, o)
    print("one-step rel L2:", errs)
    assert max(errs.values()) <= 1e-10, errs


@pytest.mark.parametrize("solver", [0, 1])
def test_trajectory_config0_replayed(O, solver):
    """Config-0 (64x64) trajectory with the oracle's dt sequence replayed:
    relative L2 per field at passes 10, 100 and 1000 of the main loop
    (Jacobi-PCG and multigrid-preconditioned PCG), bound 1e-8 after 1000."""
    from fluidsim_amd import configs

    cfg = configs.config0(rtol=1e-13, solver=solver)
    d, o = _pair(O, cfg, seed=0)
    done = nlog = 0
    for n in (10, 100, 1000):
        o.run(n - done)
        log = o.log()
        d.set_replay(log[nlog:])
        d.run(n - done)
        done, nlog = n, len(log)
        errs = _compare(d, o)
        print(f"{n}-pass rel L2:", errs)
        assert max(errs.values()) <= 1e-8, (n, errs)


def test_config1_full_size_replayed(O):
    """The metric's workload at its full size (512x512, 2 Torr, 400 V pulse):
    the two Poisson-only warm-up passes and three transport passes with the
    oracle's dt decisions replayed, both Poisson solves to a relative residual
    of 1e-13 (the device with multigrid-PCG, the oracle with Jacobi-PCG).

    phi is fixed by the stopping test only up to the system's conditioning:
    two solutions of the same system with relative residual 1e-13 differ by up
    to cond * 1e-13 in the correction (measured 5e-13 .. 5e-9 relative in phi
    here, 20x less than at 1e-12), and through the drift term that difference
    reaches the next step's densities.  So each pass starts from the oracle's
    phi: the densities then measure one pass of the device against the oracle
    (bound 1e-10, relative L2), phi gets a 1e-7 bound, and the device solve
    must meet the residual test itself."""
    import os

    from fluidsim_amd import Discharge, configs

    cfg = configs.config1(512, rtol=1e-13)
    d = Discharge(cfg)
    o = O.Oracle(oracle_cfg(cfg, solver=2, nthreads=min(os.cpu_count() or 1, 16)))
    d.init_gaussian(seed=0)
    o.init_gaussian(seed=0)
    nlog = 0
    for k in range(5):
        o.run(1)
        log = o.log()
        d.set_replay(log[nlog:])
        d.run(1)
        nlog = len(log)
        errs = _compare(d, o)
        print(f"512x512 pass {k + 1} rel L2:", errs)
        dens = {f: errs[f] for f in ("ni", "ne", "nt", "nm")}
        assert max(dens.values()) <= 1e-10, (k, errs)
        assert errs["ph"] <= 1e-7, (k, errs)
        dl = d.log()
        if k >= 2:
            assert dl[-1][6] == 0 and dl[-1][8] <= 1e-13, (k, dl[-1])
        d.set("ph", 1, o.get("ph"))


def test_config1_full_size_fifty_passes_without_resync(O):
    """The metric's workload at full size (512x512) over 52 passes (the two
    Poisson-only passes and 50 transport passes), device and oracle each
    solving its own Poisson systems to 1e-13 (multigrid-PCG / BiCGStab+ILU(0))
    with the oracle's dt decisions replayed, and phi never copied across.

    The two solutions of each Poisson system differ by up to cond * 1e-13 and
    the drift term carries that into the next step's densities; the per-field
    error curve is printed every 10 passes.  Measured on MI355X: the density
    difference peaks in the first ten transport passes (ne 6.7e-10, nt 1.1e-9)
    and then decays (1.8e-10 at pass 52), phi's peak is 1.8e-8 in an early
    pass.  Bounds: densities 1e-8 over the run (the north star's multi-step
    bound), 1e-9 at the last pass (the decay), phi 1e-7."""
    import os

    from fluidsim_amd import Discharge, configs

    cfg = configs.config1(512, rtol=1e-13)
    d = Discharge(cfg)
    o = O.Oracle(oracle_cfg(cfg, solver=1, nthreads=min(os.cpu_count() or 1, 16)))
    d.init_gaussian(seed=0)
    o.init_gaussian(seed=0)
    nlog = 0
    worst = {}
    for k in range(52):
        o.run(1)
        log = o.log()
        d.set_replay(log[nlog:])
        d.run(1)
        nlog = len(log)
        errs = _compare(d, o)
        for f, e in errs.items():
            worst[f] = max(worst.get(f, 0.0), e)
        if k % 10 == 1 or k == 51:
            print(f"512x512 pass {k + 1} rel L2:", {f: f"{e:.2e}" for f, e in errs.items()})
    print("512x512 worst over 52 passes:", {f: f"{e:.2e}" for f, e in worst.items()})
    assert max(worst[f] for f in ("ni", "ne", "nt", "nm")) <= 1e-8, worst
    assert max(errs[f] for f in ("ni", "ne", "nt", "nm")) <= 1e-9, errs
    assert worst["ph"] <= 1e-7, worst
    # the log's residual is recomputed from the final iterate after the
    # recurrence met 1e-13; that close to the rounding floor the two differ
    # (measured: up to 1.07e-13)
    dl = d.log()
    acc = dl[:, 5] == 1
    assert np.all(dl[acc, 6] == 0) and np.all(dl[acc, 8] <= 2e-13)
