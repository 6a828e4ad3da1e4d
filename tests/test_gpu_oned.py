"""GPU: the one-dimensional mode (ny = 1, the reference's default,
main.F90:29) through the C ABI against the oracle.

The stage, assembly and circuit kernels leave out every y term on a single
row (ptcl_lib.F90:285, 455; lapl_lib.F90:255, 462, 577), weight the electrode
current with w**2 pi (circ_lib.F90:98-99) and set the ghost potential to the
source voltage (circ_lib.F90:199-202); the Poisson system is tridiagonal and
solved by the Jacobi-preconditioned PCG.  Bounds as the two-dimensional
replayed trajectories'."""
import numpy as np
import pytest

from conftest import oracle_cfg, rel_l2

pytestmark = pytest.mark.gpu


def _cfg(nx, **kw):
    from fluidsim_amd import configs

    d = dict(rtol=1e-13)
    d.update(kw)
    return configs.base(nx, 1, **d)


def test_oned_grid_and_initial_state_match_oracle_bitwise(O):
    from fluidsim_amd import Discharge

    for kw in ({}, {"grid_x": 2}, {"cyl": 1, "rwall": 1}):
        c = _cfg(40, **kw)
        d = Discharge(c)
        o = O.Oracle(oracle_cfg(c))
        gd, go = d.grid(), o.grid()
        for k in go:
            assert np.array_equal(gd[k], go[k]), (kw, k)
    d.init_gaussian(seed=4)
    o.init_gaussian(seed=4)
    for name in ("ni", "ne", "nt", "nm"):
        assert np.array_equal(d.get(name, 1), o.get(name, 1)), name


def _replayed(O, c, n, seed=0):
    from fluidsim_amd import Discharge

    d = Discharge(c)
    o = O.Oracle(oracle_cfg(c))
    d.init_gaussian(seed=seed)
    o.init_gaussian(seed=seed)
    o.run(n)
    d.set_replay(o.log())
    d.run(n)
    errs = {f: rel_l2(d.get(f, 1)[1, 1:-1], o.get(f, 1)[1, 1:-1])
            for f in ("ni", "ne", "nt", "nm", "ph")}
    return d, o, errs


PULSE = dict(rf=4, p_torr=2.0, volts=400.0, t_rise=0.01, t_width=0.1, t_fall=0.01,
             ratio_x=1.05 ** (64.0 / 512), grid_x=2)


@pytest.mark.parametrize("nx,kw,n", [
    (40, dict(grid_x=0), 200),               # the reference's default grid, DC
    (40, dict(grid_x=0, rf=1), 60),          # AC: ghost potential from Vsrc, Vd from the RC stage
    (64, dict(grid_x=2, circuit=1), 60),     # active ballast, geometric grid
    (512, PULSE, 40),                        # the 512-cell trapezoid pulse
])
def test_oned_replayed_parity(O, nx, kw, n):
    d = dict(volts=300.0, res_ohm=1e5)
    d.update(kw)
    rf = d.get("rf", 0)
    c = _cfg(nx, **d)
    d, o, errs = _replayed(O, c, n)
    print(f"1-D nx={nx} {kw}, {n} passes:", errs)
    assert max(errs.values()) <= 1e-8, errs
    sd, so = d.scalars(), o.scalars()
    for k in ("Vd_pl", "Vd_mi", "Vd_or", "Vsrc", "t"):
        assert sd[k] == pytest.approx(so[k], rel=1e-12), k
    for k in ("Ida", "Idc"):
        assert so[k] != 0.0
        assert sd[k] == pytest.approx(so[k], rel=1e-8), k
    if rf == 1:
        ph = d.get("ph")
        assert ph[1, 0] == sd["Vsrc"] / 2.0 and ph[1, -1] == sd["Vsrc"] / -2.0


def test_oned_one_pass_parity(O):
    """One main-loop pass from the same state, replayed: 1e-10 (as 2-D)."""
    c = _cfg(40, volts=250.0)
    from fluidsim_amd import Discharge

    d = Discharge(c)
    o = O.Oracle(oracle_cfg(c))
    o.init_gaussian(seed=1)
    o.run(8)
    for name in ("ni", "ne", "nt", "nm"):
        for lev in (1, 2, 3):
            d.set(name, lev, o.get(name, lev))
    d.set("ph", 1, o.get("ph"))
    s = o.scalars()
    d.set_scalars({k: s[k] for k in d.SCALARS})
    o.clear_log()
    o.run(1)
    d.set_replay(o.log())
    d.run(1)
    for f in ("ni", "ne", "nt", "nm", "ph"):
        e = rel_l2(d.get(f, 1)[1, 1:-1], o.get(f, 1)[1, 1:-1])
        assert e <= 1e-10, (f, e)


def test_oned_solver_choices_and_ensemble_rejection():
    """In one dimension the multigrid choice runs the Jacobi-preconditioned
    solve (no hierarchy), the learned choice is refused, and ensembles refuse
    one-dimensional members."""
    from fluidsim_amd import Discharge, lib
    from fluidsim_amd.ensemble import Ensemble

    a = Discharge(_cfg(40, solver=1))
    b = Discharge(_cfg(40, solver=0))
    for d in (a, b):
        d.init_gaussian(seed=2)
        d.run(10)
    for f in ("ne", "ph"):
        assert np.array_equal(a.get(f, 1), b.get(f, 1)), f
    c = Discharge(_cfg(40, solver=2))
    c.init_gaussian(seed=2)
    with pytest.raises(lib.FsError):
        c.run(3)
    with pytest.raises(lib.FsError):
        Ensemble([_cfg(40), _cfg(40)])
