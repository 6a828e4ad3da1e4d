"""GPU: the stage kernel's tile heights (fs_params.stage_tile) and the timing
choice between them.

The stage kernel runs on tiles of 64 x 4, 64 x 8 or 64 x 16 cells.  Its error
sums are reduced per row segment of 64 cells with one reduction tree, so the
height changes the speed and not a bit of the result: trajectories with each
height are compared bitwise (fields, potential, wall charge, the attempt log),
on grids whose height is not a multiple of 16, with the dielectric wall and in
cylindrical geometry.  With stage_tile = 0 a handle (and an ensemble) takes
the fastest height measured on its own grid.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(cfg, passes):
    from fluidsim_amd import Discharge

    d = Discharge(cfg)
    d.init_gaussian(seed=4)
    d.run(passes)
    out = {f: d.get(f, lev) for f in ("ni", "ne", "nt", "nm") for lev in (1, 2, 3)}
    out["ph"] = d.get("ph")
    out["log"] = d.log()
    if cfg.get("rwall"):
        out["sig"] = d.get_sig("sig_pl")
    ty = d.stage_tile()
    d.close()
    return out, ty


@pytest.mark.parametrize("kw", [dict(), dict(rwall=1), dict(cyl=1)])
def test_tile_heights_give_identical_trajectories(kw):
    from fluidsim_amd import configs

    base = configs.base(100, 70, volts=300.0, rtol=1e-13, **kw)
    ref = None
    for ty in (4, 8, 16):
        out, got = _run(dict(base, stage_tile=ty), 25)
        assert got == ty
        if ref is None:
            ref = out
            continue
        for k in ref:
            assert np.array_equal(out[k], ref[k]), (ty, k)
    print("attempts", len(ref["log"]), "rejected", int(np.sum(ref["log"][:, 5] == 0)))


def test_timed_choice_and_bad_value():
    from fluidsim_amd import Discharge, configs
    from fluidsim_amd.lib import FsError

    for n in (64, 512):
        d = Discharge(configs.config1(n) if n == 512 else configs.config0())
        ty = d.stage_tile()
        print(f"{n}x{n}: stage tile 64 x {ty}")
        assert ty in (4, 8, 16)
        d.close()
    with pytest.raises(FsError):
        Discharge(configs.base(40, 30, stage_tile=5))


def test_ensemble_timed_choice_matches_fixed_heights():
    """An ensemble that times its batched launch ends bitwise where the same
    ensemble with each fixed height ends."""
    from fluidsim_amd import configs
    from fluidsim_amd.ensemble import Ensemble

    cfgs = configs.config3(3, base_seed=2, lo=64, hi=130)
    res = []
    for ty in (0, 4, 16):
        E = Ensemble([dict(c, stage_tile=ty) for c in cfgs])
        got = E.stage_tile()
        assert got in (4, 8, 16) and (ty == 0 or got == ty)
        for m, c in zip(E.members, cfgs):
            m.init_gaussian(seed=c["seed"])
        E.run(6)
        res.append([m.get("ne", 1) for m in E.members] + [m.get("ph") for m in E.members] +
                   [m.log() for m in E.members])
        print("ensemble stage tile", ty, "->", got)
        E.close()
    for other in res[1:]:
        for a, b in zip(res[0], other):
            assert np.array_equal(a, b)
