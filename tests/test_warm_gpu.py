"""GPU: the multigrid-PCG warm start (fs_params.lapl_warm).  A solve after an
accepted transport step starts from the previous converged correction; the
trajectory stays within the solver tolerance of the cold-start one (same dt
decisions replayed), the solves need fewer iterations, and a field change
(fs_set_field) or a solve without a transport step in between starts from
zero again."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _iters(log):
    return [int(r[7]) for r in log if r[6] >= 0]


def test_warm_start_follows_the_cold_trajectory():
    from fluidsim_amd import Discharge, configs

    passes = 40
    cold = Discharge(configs.config0(rtol=1e-12))
    cold.init_gaussian(seed=0)
    cold.run(passes)
    warm = Discharge(configs.config0(rtol=1e-12, warm=1))
    warm.init_gaussian(seed=0)
    warm.set_replay(cold.log())
    warm.run(passes)
    errs = {f: rel_l2(warm.get(f, 1), cold.get(f, 1)) for f in ("ni", "ne", "nt", "nm")}
    errs["ph"] = rel_l2(warm.get("ph"), cold.get("ph"))
    ic, iw = _iters(cold.log()), _iters(warm.log())
    print("rel L2 warm vs cold after", passes, "passes:", errs)
    print("iterations cold", ic)
    print("iterations warm", iw)
    assert max(errs[f] for f in ("ni", "ne", "nt", "nm")) <= 1e-10, errs
    assert errs["ph"] <= 1e-9, errs
    assert all(r[6] == 0 for r in warm.log() if r[6] >= 0)
    # the first passes are Poisson-only (no warm start); afterwards fewer iterations
    assert sum(iw[5:]) < sum(ic[5:]), (ic, iw)


def test_field_change_and_repeated_solves_start_from_zero():
    from fluidsim_amd import Discharge, configs

    d = Discharge(configs.config0(warm=1))
    d.init_gaussian(seed=0)
    d.run(8)
    ph = d.get("ph")
    sols, its = [], []
    for _ in range(2):
        d.set("ph", 1, ph)  # resets the warm start
        lerr, st = d.lapl_solve()
        assert lerr == 0
        sols.append(d.get("ph"))
        its.append(st.iters)
    assert its[0] == its[1] and np.array_equal(sols[0], sols[1]), its
    # a solve right after a solve (no transport step in between) also starts
    # from zero: it equals the same solve made cold by a field reset
    lerr, st_rep = d.lapl_solve()
    sol_rep = d.get("ph")
    d.set("ph", 1, sols[1])
    lerr_c, st_c = d.lapl_solve()
    assert lerr == 0 and lerr_c == 0
    assert st_rep.iters == st_c.iters and np.array_equal(sol_rep, d.get("ph"))


@pytest.mark.parametrize("cfg_name", ["config0", "config2_256"])
def test_projected_warm_start_follows_the_cold_trajectory(cfg_name):
    """lapl_warm = 2: each solve starts from the energy-norm projection onto the
    last three converged corrections.  The trajectory stays within the solver
    tolerance of the cold-start one (dt decisions replayed) and the solves need
    fewer iterations; on the 2048^2 workload's geometry (here at 256^2) the
    projection also needs no more iterations than the extrapolation."""
    from fluidsim_amd import Discharge, configs

    def mk(warm):
        if cfg_name == "config0":
            return configs.config0(rtol=1e-12, warm=warm)
        return configs.config2(256, rtol=1e-12, warm=warm)

    passes = 40
    cold = Discharge(mk(0))
    cold.init_gaussian(seed=0)
    cold.run(passes)
    runs = {}
    for w in (1, 2):
        d = Discharge(mk(w))
        d.init_gaussian(seed=0)
        d.set_replay(cold.log())
        d.run(passes)
        runs[w] = d
    d = runs[2]
    errs = {f: rel_l2(d.get(f, 1), cold.get(f, 1)) for f in ("ni", "ne", "nt", "nm")}
    errs["ph"] = rel_l2(d.get("ph"), cold.get("ph"))
    ic, i1, i2 = _iters(cold.log()), _iters(runs[1].log()), _iters(d.log())
    print("rel L2 projected vs cold after", passes, "passes:", errs)
    print("iterations cold", ic)
    print("iterations extrapolated", i1)
    print("iterations projected", i2)
    assert max(errs[f] for f in ("ni", "ne", "nt", "nm")) <= 1e-10, errs
    assert errs["ph"] <= 1e-9, errs
    assert all(r[6] == 0 for r in d.log() if r[6] >= 0)
    assert sum(i2[5:]) < sum(ic[5:]), (ic, i2)
    if cfg_name != "config0":
        assert sum(i2[5:]) <= sum(i1[5:]), (i1, i2)


def test_warm_start_argument_checks():
    """lapl_warm takes 0, 1 or 2, for single handles and ensemble members."""
    from fluidsim_amd import Discharge, configs
    from fluidsim_amd.ensemble import Ensemble
    from fluidsim_amd.lib import FsError

    with pytest.raises(FsError):
        Discharge(configs.config0(warm=3))
    cfgs = configs.config3(2, base_seed=1, lo=64, hi=72)
    for c in cfgs:
        c["lapl_warm"] = 3
    with pytest.raises(FsError):
        Ensemble(cfgs)
