"""GPU: grid, initial state, failure paths and model variants through the C ABI.

Grid metrics and the Gaussian initial state are computed on the host by the
library and must equal the oracle's bitwise.  Failure paths are forced from
the host (negative density, dt floor, a solve capped below convergence,
an exhausted replay table), never by making a kernel fault.
"""
import numpy as np
import pytest

from conftest import oracle_cfg, rel_l2

pytestmark = pytest.mark.gpu


def _cfg(nx=24, ny=16, **kw):
    from fluidsim_amd import configs

    return configs.base(nx, ny, rtol=1e-13, **kw)


@pytest.mark.parametrize("gx,gy,cyl", [(0, 0, 0), (1, 1, 0), (2, 0, 0), (2, 2, 1)])
def test_grid_matches_oracle_bitwise(O, gx, gy, cyl):
    from fluidsim_amd import Discharge

    c = _cfg(30, 20, grid_x=gx, cyl=cyl)
    c["grid_y"], c["ratio_y"] = gy, 1.03
    d = Discharge(c)
    o = O.Oracle(oracle_cfg(c))
    gd, go = d.grid(), o.grid()
    for k in go:
        assert np.array_equal(gd[k], go[k]), k


def test_gaussian_initial_state_matches_oracle_bitwise(O):
    from fluidsim_amd import Discharge

    c = _cfg(20, 14)
    d = Discharge(c)
    o = O.Oracle(oracle_cfg(c))
    d.init_gaussian(seed=11, noise=0.05)
    o.init_gaussian(seed=11, noise=0.05)
    for name in ("ni", "ne", "nt", "nm"):
        for lev in (1, 2, 3):
            assert np.array_equal(d.get(name, lev), o.get(name, lev)), (name, lev)
    d.ptcl_init()
    o.init_uniform()
    assert np.array_equal(d.get("nt", 2), o.get("nt", 2))


def test_level_writes_keep_levels_apart():
    """After an accept the device holds levels 1 and 2 in one buffer; a write
    of either level gives level 1 its own copy first, so the other keeps its
    values (a host that rotates its arrays, as the reference's main loop does,
    writes them one at a time)."""
    from fluidsim_amd import Discharge

    d = Discharge(_cfg(20, 14))
    d.init_gaussian(seed=5)
    d.run(3)
    for name in ("ne", "nt"):
        l1, l3 = d.get(name, 1), d.get(name, 3)
        x = l1 * 1.5
        d.set(name, 2, x)
        assert np.array_equal(d.get(name, 2), x)
        assert np.array_equal(d.get(name, 1), l1)
        assert np.array_equal(d.get(name, 3), l3)
    d.run(2)
    d.set("ni", 1, d.get("ni", 1) * 0.5)
    assert not np.array_equal(d.get("ni", 1), d.get("ni", 2))


def test_circ_init_sets_amplitude_and_voltages():
    """circ_init(Vset) (circ_lib.F90:213-223) on a handle created with
    vset = 0 equals creating it with vset."""
    from fluidsim_amd import Discharge

    c = _cfg(20, 14, volts=250.0)
    a = Discharge(c)
    b = Discharge(dict(c, vset=0.0))
    b.circ_init(c["vset"])
    sb = b.scalars()
    for k in ("Vd_pl", "Vd_mi", "Vd_or", "Vsrc"):
        assert sb[k] == c["vset"], k
    for d in (a, b):
        d.init_gaussian(seed=2)
        d.run(4)
    assert np.array_equal(a.get("ne", 1), b.get("ne", 1))
    assert np.array_equal(a.get("ph"), b.get("ph"))
    assert a.scalars()["Vd_pl"] == b.scalars()["Vd_pl"]


def test_negative_density_is_fatal(O):
    from fluidsim_amd import Discharge, lib

    d = Discharge(_cfg())
    d.init_gaussian(seed=1)
    d.run(3)
    nt = d.get("nt", 2)
    nt[4, 5] = -1.0
    d.set("nt", 2, nt)
    with pytest.raises(lib.FsError) as e:
        d.ptcl_step()
    assert e.value.code == -2


def test_dt_floor_is_fatal(O):
    from fluidsim_amd import Discharge, lib

    d = Discharge(_cfg())
    d.init_gaussian(seed=1)
    d.run(2)
    d.set_scalars({"dt": 4e-13})
    with pytest.raises(lib.FsError) as e:
        d.ptcl_step()
    assert e.value.code == -4


def test_unconverged_solve_takes_the_reject_path(O):
    from fluidsim_amd import Discharge

    c = _cfg()
    c["lapl_rtol"], c["lapl_maxit"] = 1e-30, 3
    d = Discharge(c)
    d.init_gaussian(seed=1)
    s0 = d.scalars()
    ph0, n3 = d.get("ph"), d.get("ne", 3)
    lerr, st = d.step()
    s1 = d.scalars()
    assert lerr == 1 and st.iters >= 3
    assert s1["dt"] == s0["dt"] / 2 and s1["ts"] == 0
    assert np.array_equal(d.get("ph"), ph0)
    assert np.array_equal(d.get("ne", 1), n3) and np.array_equal(d.get("ne", 2), n3)


def test_exhausted_replay_is_fatal(O):
    from fluidsim_amd import Discharge, lib

    c = _cfg()
    o = O.Oracle(oracle_cfg(c))
    d = Discharge(c)
    o.init_gaussian(seed=2)
    d.init_gaussian(seed=2)
    o.run(3)
    d.set_replay(o.log()[:1])
    with pytest.raises(lib.FsError) as e:
        d.run(6)
    assert e.value.code == -7


def _replayed_parity(O, c, n, seed=0, fields=("ni", "ne", "nt", "nm", "ph"), init=None):
    from fluidsim_amd import Discharge

    d = Discharge(c)
    o = O.Oracle(oracle_cfg(c))
    d.init_gaussian(seed=seed)
    o.init_gaussian(seed=seed)
    if init:
        init(d, o)
    o.run(n)
    d.set_replay(o.log())
    d.run(n)
    errs = {}
    for f in fields:
        a = d.get(f, 1)[1:-1, 1:-1]
        b = o.get(f, 1)[1:-1, 1:-1]
        errs[f] = rel_l2(a, b)
    return d, o, errs


def test_cylindrical_geometry_parity(O):
    c = _cfg(24, 20, cyl=1, grid_x=0, ratio_x=1.0)
    d, o, errs = _replayed_parity(O, c, 12)
    print("cyl:", errs)
    assert max(errs.values()) <= 1e-10, errs


def test_partial_electrode_parity(O):
    c = _cfg(24, 20)
    c["ew"], c["gap"] = 0.4 * c["w"], 0.8 * c["w"]
    d, o, errs = _replayed_parity(O, c, 12)
    print("partial electrode:", errs)
    assert max(errs.values()) <= 1e-10, errs
    sd, so = d.scalars(), o.scalars()
    for k in ("Vd_pl", "Ida", "Idc"):
        assert sd[k] == pytest.approx(so[k], rel=1e-9, abs=1e-12 * abs(so["Vd_pl"])), k


@pytest.mark.parametrize("rf,circuit", [(0, 1), (1, 0), (4, 0)])
def test_circuit_variants_parity(O, rf, circuit):
    """Active RC ballast (DC), AC source through the reference's RC stage, and
    the trapezoid pulse: fields and circuit scalars after replayed steps."""
    c = _cfg(24, 16, rf=rf, circuit=circuit, volts=300.0, res_ohm=1e5,
             t_rise=1e-6, t_width=1e-5, t_fall=1e-6)
    d, o, errs = _replayed_parity(O, c, 14)
    print(f"rf={rf} circuit={circuit}:", errs)
    assert max(errs.values()) <= 1e-10, errs
    sd, so = d.scalars(), o.scalars()
    for k in ("Vd_pl", "Vd_mi", "Vd_or", "Vsrc"):
        assert sd[k] == pytest.approx(so[k], rel=1e-12), k
    for k in ("Ida", "Idc"):
        assert sd[k] == pytest.approx(so[k], rel=1e-8, abs=1e-14), k


def test_graph_replay_equals_eager_launches(O):
    """The MG-PCG iterations replayed from a captured HIP graph give bitwise the
    same fields as the same kernels launched one by one, over many solves."""
    from fluidsim_amd import Discharge, configs

    out = []
    for graphs in (0, 1):
        c = configs.base(40, 28, rtol=1e-12, solver=1, graphs=graphs)
        d = Discharge(c)
        d.init_gaussian(seed=5)
        d.run(16)
        out.append({k: d.get(k, 1) for k in ("ne", "ni", "nt", "nm")} | {"ph": d.get("ph")})
        out[-1]["log"] = d.log()
    for k in ("ne", "ni", "nt", "nm", "ph"):
        assert np.array_equal(out[0][k], out[1][k]), k
    assert np.array_equal(out[0]["log"], out[1]["log"])


def test_tiny_grid_uses_jacobi_fallback(O):
    """At <= 16 cells the hierarchy has one level; the solve must still work."""
    c = _cfg(4, 3)
    c["lapl_solver"] = 1
    d, o, errs = _replayed_parity(O, c, 8)
    assert max(errs.values()) <= 1e-10, errs


def test_small_grid_two_level_hierarchy(O):
    """17..64 cells: a two-level hierarchy whose coarse level is the dense solve."""
    c = _cfg(8, 6)
    c["lapl_solver"] = 1
    d, o, errs = _replayed_parity(O, c, 8)
    assert max(errs.values()) <= 1e-10, errs


def test_indefinite_operator_solved_by_multigrid_pcg(O):
    """dFdPh (lapl_lib.F90:611-636) subtracts two terms of size dn/a at a small
    SG argument a; at a near-zero field across a density step the difference
    is rounding noise and can leave a face coupling of the wrong sign, so the
    assembled operator is symmetric but indefinite.  A potential step of 1e-6
    across one face with a 5x density step gives one such face (and, in the
    cell-volume-weighted form, one negative eigenvalue).  The multigrid-PCG
    solve must step along the negative-curvature direction and converge in a
    multigrid iteration count (not stop, and not fall back to Jacobi-PCG), to
    the direct solution of the assembled system (the CPU oracle's banded direct
    solve of this state agrees with that solution to 1e-14)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    from fluidsim_amd import Discharge, configs

    c = configs.config1(64, rtol=1e-10, debug=1, warm=0)
    c["lapl_solver"] = 1
    d = Discharge(c)
    d.init_gaussian(seed=0)
    d.ptcl_step()
    g = d.grid()
    jc, ic = 31, 31  # the face between interior rows jc and jc + 1 at column ic
    ph = g["y"][:, None] * np.ones_like(d.get("ph"))
    ph[jc + 2, ic + 1] = ph[jc + 1, ic + 1] + 1e-6
    ni, ne = d.get("ni", 1), d.get("ne", 1)
    ni[jc + 2, ic + 1] *= 5.0
    ne[jc + 2, ic + 1] *= 5.0
    d.set("ph", 1, ph)
    d.set("ni", 1, ni)
    d.set("ne", 1, ne)
    b, A5 = d.lapl_assemble()

    ny, nx = b.shape
    assert np.argwhere(A5[..., :4] < 0).tolist() == [[jc, ic, 3], [jc + 1, ic, 2]]
    assert (A5[..., 4] < 0).all()
    idx = np.arange(nx * ny).reshape(ny, nx)
    rows, cols, vals = [idx.ravel()], [idx.ravel()], [A5[..., 4].ravel()]
    for k, (sl, nb) in enumerate((((slice(None), slice(1, None)), (slice(None), slice(None, -1))),
                                  ((slice(None), slice(None, -1)), (slice(None), slice(1, None))),
                                  ((slice(1, None), slice(None)), (slice(None, -1), slice(None))),
                                  ((slice(None, -1), slice(None)), (slice(1, None), slice(None))))):
        rows.append(idx[sl].ravel())
        cols.append(idx[nb].ravel())
        vals.append(A5[..., k][sl].ravel())
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(nx * ny,) * 2)
    V = sp.diags(np.outer(g["dly"], g["dlx"]).ravel())
    M = -(V @ A)
    assert abs(M - M.T).max() <= 1e-12 * abs(M).max()
    ev = np.linalg.eigvalsh(M.toarray())
    assert ev[0] < 0 < ev[1], ev[:3]
    # the solve is for the correction: A sol = -b, then ph += sol (lapl_lib.F90:89-119)
    x = spla.spsolve(A.tocsc(), -b.ravel()).reshape(ny, nx)

    lerr, st = d.lapl_solve()
    assert lerr == 0
    assert st.iters <= 40, st.iters
    e = rel_l2(d.get("ph")[1:-1, 1 : nx + 1] - ph[1:-1, 1 : nx + 1], x)
    assert e <= 1e-8, e


def test_run_issuing_passes_ahead_equals_single_passes(O):
    """Within one fs_run call the next pass's accept, begin and first stage are
    issued behind each solve's first chunk, gated on the device by that pass's
    outcome; one pass per call issues nothing ahead.  Both must give the same
    fields, scalars and log bit for bit, over passes whose solves need more
    than their first chunk (the gate closed) and passes that need only it."""
    from fluidsim_amd import Discharge, configs

    c = configs.config1(128, rtol=1e-10, chunk=32, warm=1)
    a, b = Discharge(c), Discharge(c)
    a.init_gaussian(seed=3)
    b.init_gaussian(seed=3)
    a.run(40)
    for _ in range(40):
        b.run(1)
    for f in ("ni", "ne", "nt", "nm", "ph"):
        assert np.array_equal(a.get(f, 1), b.get(f, 1)), f
    assert a.scalars() == b.scalars()
    la, lb = a.log(), b.log()
    assert np.array_equal(la, lb)
    iters = la[la[:, 6] >= 0, 7]
    assert len(set(iters.tolist())) > 1, iters  # first chunks both long enough and too short
