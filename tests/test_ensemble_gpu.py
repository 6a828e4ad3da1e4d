"""GPU: ensembles.  Members advanced together by the ensemble's batched
launches end bitwise where the same discharges end when run alone (fields,
potential, scalars and the attempt log), across different grid sizes,
pressures, voltages and secondary-emission coefficients."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _solo(c, passes):
    from fluidsim_amd import Discharge

    d = Discharge(c)
    d.init_gaussian(seed=c["seed"])
    d.run(passes)
    return d


@pytest.mark.parametrize("passes,warm,n,lo,hi", [(1, 0, 5, 64, 96), (14, 0, 5, 64, 96),
                                                 (14, 1, 5, 64, 96), (6, 1, 2, 256, 300),
                                                 (14, 2, 5, 64, 96), (6, 2, 2, 256, 300)])
def test_members_equal_their_solo_runs_bitwise(passes, warm, n, lo, hi):
    """(256..300 cells a side: the assembly coarsens three multigrid levels
    itself in a solo run, while the ensemble's coarsening launches rewrite
    them, so equal bits also show the two coarsenings agree.)"""
    from fluidsim_amd import configs
    from fluidsim_amd.ensemble import Ensemble

    cfgs = configs.config3(n, base_seed=3, lo=lo, hi=hi)
    for c in cfgs:
        c["lapl_warm"] = warm
    E = Ensemble(cfgs)
    for m, c in zip(E.members, cfgs):
        m.init_gaussian(seed=c["seed"])
    E.run(passes)
    for i, (m, c) in enumerate(zip(E.members, cfgs)):
        d = _solo(c, passes)
        for f in ("ni", "ne", "nt", "nm"):
            for lev in (1, 2, 3):
                assert np.array_equal(m.get(f, lev), d.get(f, lev)), (i, f, lev)
        assert np.array_equal(m.get("ph"), d.get("ph")), i
        assert np.array_equal(m.log(), d.log()), i
        assert m.scalars() == d.scalars(), i
        print(i, (c["nx"], c["ny"]), "attempts", len(d.log()), "t", d.scalars()["t"])
    E.close()


def test_ensemble_runs_continue_like_one_run():
    """Two calls of run(k) equal one call of run(2k), member by member."""
    from fluidsim_amd import configs
    from fluidsim_amd.ensemble import Ensemble

    cfgs = configs.config3(3, base_seed=9, lo=64, hi=80)
    out = []
    for split in (False, True):
        E = Ensemble(cfgs)
        for m, c in zip(E.members, cfgs):
            m.init_gaussian(seed=c["seed"])
        if split:
            E.run(5)
            E.run(5)
        else:
            E.run(10)
        out.append([m.get("ne", 1) for m in E.members] + [m.get("ph") for m in E.members])
        E.close()
    for a, b in zip(*out):
        assert np.array_equal(a, b)


def test_graph_rounds_equal_eager_rounds():
    """The rounds replayed from HIP graphs (members with use_graphs) equal the
    same rounds launched one kernel at a time, bitwise, over two runs (each
    run captures its graphs anew) with rejected attempts and multi-chunk
    solves among them."""
    from fluidsim_amd import configs
    from fluidsim_amd.ensemble import Ensemble

    cfgs = configs.config3(4, base_seed=5, lo=64, hi=128)
    out = []
    for graphs in (0, 1):
        E = Ensemble([dict(c, use_graphs=graphs) for c in cfgs])
        for m, c in zip(E.members, cfgs):
            m.init_gaussian(seed=c["seed"])
        E.run(6)
        E.run(5)
        out.append(([m.get(f, 1) for m in E.members for f in ("ni", "ne", "nt", "nm")],
                    [m.get("ph") for m in E.members], [m.log() for m in E.members]))
        E.close()
    (fa, pa, la), (fb, pb, lb) = out
    for a, b in zip(fa + pa + la, fb + pb + lb):
        assert np.array_equal(a, b)
    print("attempts per member", [len(x) for x in lb],
          "rejected", [int(np.sum(x[:, 5] == 0)) for x in lb])
