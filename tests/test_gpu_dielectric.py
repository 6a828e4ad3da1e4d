"""GPU: the dielectric wall (rwall = 1, top row type_y = 3) against the oracle.

* the grid's wall row types, device against oracle;
* one RHS sweep per stage and one Poisson assembly from an identical state
  with a nonzero, uneven wall charge (sig_mi for the transport wall face and
  the Joule term, sig_pl for the Poisson row), relative L2 1e-12;
* the main loop with the oracle's dt decisions replayed, including rejected
  attempts (the charge's restore) and accepted ones (its rotation), with and
  without HIP graphs: densities, phi and the three charge levels at passes
  12, 32 and 62, bound 1e-8 (the north star's multi-step bound).
"""
import numpy as np
import pytest

from conftest import oracle_cfg, rel_l2

pytestmark = pytest.mark.gpu

SIG = ("sig_pl", "sig_mi", "sig_or")


def _cfg(**kw):
    from fluidsim_amd import configs

    return configs.base(48, 40, rwall=1, volts=300.0, rtol=1e-13, **kw)


def test_wall_row_types_match_oracle(O):
    from fluidsim_amd import Discharge

    cfg = _cfg()
    d = Discharge(cfg)
    o = O.Oracle(oracle_cfg(cfg))
    gd, go = d.grid(), o.grid()
    assert np.array_equal(gd["type_y"], go["type_y"])
    assert np.all(gd["type_y"][-1] == 3)
    d.close()


def test_rhs_sweep_and_assembly_with_wall_charge(O):
    from fluidsim_amd import Discharge

    cfg = _cfg(debug=1)
    d = Discharge(cfg)
    o = O.Oracle(oracle_cfg(cfg))
    d.init_gaussian(seed=1)
    o.init_gaussian(seed=1)
    # a potential with gradients everywhere, then a wall charge of the size of
    # the applied field with sign changes along the wall
    ny, P = d.ny + 2, d.P
    yy, xx = np.meshgrid(np.linspace(0, 1, ny), np.linspace(0, 1, P), indexing="ij")
    ph = cfg["vset"] * (0.5 - xx + 0.1 * np.sin(3.0 * yy + 2.0 * xx))
    d.set("ph", 1, ph)
    o.set("ph", 1, ph)
    e0 = cfg["vset"] / cfg["l"]
    xs = np.linspace(0, 1, d.nx + 2)
    rng = np.random.default_rng(3)
    for which, f in (("sig_pl", 0.7), ("sig_mi", -0.4), ("sig_or", 0.2)):
        s = e0 * (f * np.cos(5.0 * xs) + 0.05 * rng.standard_normal(xs.size))
        d.set_sig(which, s)
        o.set_sig(which, s)
        assert np.array_equal(d.get_sig(which), s)
    for stage in (1, 3, 5):
        d.rhs_sweep(stage)
        o.rhs_sweep(stage)
        for name in ("ki", "ke", "kt", "km"):
            a, b = d.get(name, stage)[1:-1, 1:-1], o.get(name, stage)[1:-1, 1:-1]
            e = rel_l2(a, b)
            ew = rel_l2(a[-1], b[-1])                 # the wall row alone
            print(f"stage {stage} {name}: rel L2 {e:.2e}, wall row {ew:.2e}")
            assert e < 1e-12 and ew < 1e-12, (stage, name, e, ew)
    b_d, A_d = d.lapl_assemble()
    st, b_o, A_o = o.lapl_assemble()
    assert st == 0
    eb = rel_l2(b_d, b_o)
    ebw = rel_l2(b_d[-1], b_o[-1])
    print("assembly: b", eb, "wall row", ebw, "A", [rel_l2(A_d[..., c], A_o[..., c])
                                                     for c in range(5)])
    assert eb < 1e-12 and ebw < 1e-12
    for c in range(5):
        assert rel_l2(A_d[..., c], A_o[..., c]) < 1e-12, c
    # the wall charge is an input of the sweep and the assembly, not changed
    # by them
    for which in SIG:
        assert np.array_equal(d.get_sig(which), o.get_sig(which)), which
    d.close()


@pytest.mark.parametrize("graphs", [0, 1])
def test_trajectory_with_wall_replayed(O, graphs):
    import os

    from fluidsim_amd import Discharge

    cfg = _cfg(graphs=graphs)
    d = Discharge(cfg)
    o = O.Oracle(oracle_cfg(cfg, solver=2, nthreads=min(os.cpu_count() or 1, 16)))
    d.init_gaussian(seed=0)
    o.init_gaussian(seed=0)
    done = nlog = 0
    for n in (12, 32, 62):
        o.run(n - done)
        log = o.log()
        d.set_replay(log[nlog:])
        d.run(n - done)
        done, nlog = n, len(log)
        errs = {f: rel_l2(d.get(f, 1)[1:-1, 1:-1], o.get(f, 1)[1:-1, 1:-1])
                for f in ("ni", "ne", "nt", "nm")}
        errs["ph"] = rel_l2(d.get("ph")[1:-1, 1:-1], o.get("ph")[1:-1, 1:-1])
        for which in SIG:
            errs[which] = rel_l2(d.get_sig(which)[1:-1], o.get_sig(which)[1:-1])
        print(f"graphs {graphs} pass {n} rel L2:", {k: f"{v:.2e}" for k, v in errs.items()})
        assert max(errs.values()) <= 1e-8, (n, errs)
    sig = o.get_sig("sig_pl")
    assert np.abs(sig).max() > 0                       # the wall charged
    dl, ol = d.log(), o.log()
    assert np.any(ol[:, 5] == 0)                       # rejected attempts were replayed
    assert np.array_equal(dl[:, 5], ol[:, 5])
    acc = dl[:, 5] == 1
    assert np.all(dl[acc, 6] == 0) and np.all(dl[acc, 8] <= 1e-13)
    d.close()
