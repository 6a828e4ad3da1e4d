"""Pins of the CPU oracle (oracle/fsim_oracle.c) -- CPU only.

The reference ships no tests, golden vectors or fixtures, and cannot be built
here (it needs MPI and PETSc), so each oracle function is pinned by a check
that a plausible mistake (dropped term, wrong sign or index, transposed
operand, unwidened literal) would fail:

  constants         mechanical evaluation of the reference's parameter text
                    (tests/golden/fit_vectors.npz) -- bitwise; amdflang probe
                    values for the float32-widened x0 and log(5.119)
  fits (R4-R6)      mechanical evaluation of the reference's fit source text
                    (golden vectors); the reference's own double-literal Python
                    restatements of De, Dt (coeftest.py) and nu (wvgRes_data.py);
                    clamps and branch breakpoints
  get_Flux (R8)     Bernoulli form in 50-digit arithmetic, E->0 Fick limit,
                    drift limits, the -100 exponent clamp
  dFdPh (R16)       50-digit central difference of the SG flux in v
  rk_step (R11)     y' = lambda*y: Merson's stability polynomial and its
                    embedded error estimate z^5/720, in exact rationals
  controller (R12)  the recorded attempt log against the controller law
  grid (R1)         closed forms of the uniform grid; geometric ratios; types
  ptclEqn (R10)     uniform state: interior slope equals the sources only and
                    the electrode cells differ by the thermal-flux terms;
                    y-invariance of x-only states; telescoping of interior faces
  laplEqn/laplJ     laplJ = finite-difference Jacobian of laplEqn; dt = 0
  (R15, R16)        gives the nonuniform 5-point Laplacian; V*A symmetric
  solvers (R17)     banded LU = dense solve; PCG and BiCGStab+ILU(0) = LU
  main loop (R18)   reject path restores the levels and halves dt; fatal paths

Parity unpinned (no reference output exists for it): the end-to-end
trajectory of the full model against the actual Fortran program.
"""
import fractions
import os

import numpy as np
import pytest

from conftest import oracle_cfg

GOLD = os.path.join(os.path.dirname(__file__), "golden", "fit_vectors.npz")
FIT_NAMES = ["get_mue", "get_mut", "get_De", "get_Dt", "get_k_ex", "get_k_ir", "get_nu",
             "get_k_sc", "get_k_si", "get_beta"]


def small_cfg(nx=12, ny=8, **kw):
    from fluidsim_amd import configs

    c = configs.base(nx, ny, **kw)
    return c


# ---------------------------------------------------------------------------
def test_constants_match_reference_text(O):
    g = np.load(GOLD)
    k = O.consts(float(g["p_torr"]), 0.1)
    names = {"eps0": "eps0", "mu0": "mu0", "e": "e", "kb": "kb", "me": "me", "x0": "x0",
             "ph0": "ph0", "t0": "t0", "wp02": "wp02", "w2": "w2", "tg": "Tg", "p": "p",
             "ninf": "ninf", "n_zero": "n_zero", "mi": "mi", "mui": "mui", "vi": "vi",
             "di": "Di", "dm": "Dm", "k_r": "k_r", "k_mp": "k_mp", "k_2q": "k_2q",
             "h_ir": "H_ir", "h_ex": "H_ex", "h_si": "H_si", "h_sc": "H_sc", "gam": "gam"}
    for gk, ok in names.items():
        assert float(g["c_" + gk]) == k[ok], (gk, float(g["c_" + gk]), k[ok])
    # x0**6: the oracle rounds once (as gfortran folds it); the golden evaluator
    # multiplies six times, so allow 2 ulp
    assert abs(k["k_3q"] - float(g["c_k_3q"])) <= 2 * np.spacing(k["k_3q"])


def test_float32_widening_probe_values(O):
    # values printed by amdflang for the reference's literals
    k = O.consts(10.0)
    assert k["x0"] == 9.99999974737875164e-05
    assert k["log5119"] == 1.63295912742614746
    assert k["log5119"] != np.log(5.119)  # float32 log, not the double one


def test_fits_match_golden_vectors(O):
    g = np.load(GOLD)
    out = O.eval_fits(g["T"], float(g["p_torr"]))
    for c, name in enumerate(FIT_NAMES):
        rel = np.abs(out[:, c] - g[name]) / np.abs(g[name])
        assert rel.max() < 1e-11, (name, rel.max())


def test_fits_match_reference_python_restatements(O):
    """coeftest.py / wvgRes_data.py use double literals where the Fortran has
    float32-widened ones: agreement to ~1e-4 pins the coefficient-to-power map."""
    g = np.load(GOLD)
    k = O.consts(10.0)
    out = O.eval_fits(g["T"], 10.0)
    sD = k["x0"] / k["ninf"] * k["t0"]
    assert np.max(np.abs(out[:, 2] / sD / g["py_De"] - 1)) < 5e-4
    assert np.max(np.abs(out[:, 3] / sD / g["py_Dt"] - 1)) < 5e-4
    snu = k["ninf"] * k["t0"] / k["x0_3"]
    assert np.max(np.abs(out[:, 6] / snu / g["py_nu"] - 1)) < 5e-5
    # and float32 widening is visible: not equal to the double-literal forms
    assert np.max(np.abs(out[:, 2] / sD / g["py_De"] - 1)) > 1e-9


def test_fit_clamps_and_breakpoints(O):
    k = O.consts(10.0)
    ph0 = k["ph0"]
    # points strictly inside each clamped range (T*ph0 rounds, so not on the bound)
    T = np.array([0.01, 0.1, 0.2, 160.0, 1000.0, 0.9, 0.95, 1.0, 1.3, 0.3, 0.45, 17.0,
                  30.0]) / ph0
    f = O.eval_fits(T, 10.0)
    # mue/De/nu are constant below 0.234 V and above 157 V
    assert f[0, 0] == f[1, 0] == f[2, 0] and f[3, 0] == f[4, 0]
    assert f[0, 2] == f[2, 2] and f[3, 6] == f[4, 6]
    # k_ex clamps at 0.9611, k_ir at 1.37, k_sc/k_si at 0.5 and 16
    assert f[5, 4] == f[6, 4] != f[7, 4]
    assert f[7, 5] == f[8, 5]
    assert f[9, 7] == f[10, 7] and f[11, 7] == f[12, 7]
    assert f[9, 8] == f[10, 8] and f[11, 8] == f[12, 8]
    # branch switch of get_mue at the float32 log(5.119): continuous to ~1e-2
    xb = k["log5119"]
    lo, hi = np.nextafter(xb, -np.inf), xb
    Tb = np.exp(np.array([lo, hi])) / ph0
    fb = O.eval_fits(Tb, 10.0)
    assert abs(fb[0, 0] / fb[1, 0] - 1) < 0.05


# ---------------------------------------------------------------------------
def _sg_exact(E, h, q, mu, D, nl, nr):
    import mpmath as mp

    mp.mp.dps = 50
    v = mp.mpf(q) * mp.mpf(mu) * mp.mpf(E)
    a = v * mp.mpf(h) / mp.mpf(D)
    if a == 0:
        return mp.mpf(D) * (mp.mpf(nl) - mp.mpf(nr)) / mp.mpf(h)

    def B(x):
        return x / mp.expm1(x)

    return mp.mpf(D) / mp.mpf(h) * (B(-a) * mp.mpf(nl) - B(a) * mp.mpf(nr))


def test_sg_flux_bernoulli_form(O):
    rng = np.random.default_rng(3)
    n = 600
    E = rng.standard_normal(n) * 10.0 ** rng.uniform(-6, 3, n)
    h = rng.uniform(0.3, 3.0, n)
    q = rng.choice([-1.0, 1.0], n)
    mu = 10.0 ** rng.uniform(-3, 2, n)
    D = 10.0 ** rng.uniform(-2, 3, n)
    nl = 10.0 ** rng.uniform(-4, 6, n)
    nr = 10.0 ** rng.uniform(-4, 6, n)
    args = np.stack([E, h, q, mu, D, nl, nr], 1)
    got = O.eval_flux(args)
    for m in range(n):
        a = abs(q[m] * mu[m] * E[m] * h[m] / D[m])
        if a > 90:  # the -100 clamp is checked separately
            continue
        ex = float(_sg_exact(*args[m]))
        scale = abs(q[m] * mu[m] * E[m]) * (nl[m] + nr[m]) + D[m] * (nl[m] + nr[m]) / h[m]
        assert abs(got[m] - ex) <= (1e-14 + 4e-16 / min(a, 1.0)) * scale, m


def test_sg_flux_limits(O):
    # E -> 0 (|qE| < 1e-12): Fick; strong drift: upwind; beyond -100: clamped exponent
    D, h, mu = 2.0, 0.5, 3.0
    fick = O.eval_flux([[1e-13, h, 1, mu, D, 5.0, 2.0]])[0]
    assert fick == D * (5.0 - 2.0) / h
    up = O.eval_flux([[50.0, h, 1, mu, D, 5.0, 2.0], [-50.0, h, 1, mu, D, 5.0, 2.0]])
    assert abs(up[0] - mu * 50.0 * 5.0) / (mu * 50.0 * 5.0) < 1e-12
    assert abs(up[1] - (-mu * 50.0 * 2.0)) / (mu * 50.0 * 2.0) < 1e-12
    big = O.eval_flux([[1e4, h, 1, mu, D, 5.0, 2.0]])[0]
    v = mu * 1e4
    assert big == v * (5.0 - 2.0 * np.exp(-100.0)) / (1.0 - np.exp(-100.0))


def test_sg_degenerate_argument_takes_the_limit(O):
    """|qE| and |v| pass the reference's 1e-12 tests but |arg| < 2^-53, so
    1 - exp(-|arg|) is exactly 0: the reference divides by zero there and
    stops; the oracle (like the kernels) takes the a -> 0 limits, Fick's flux
    D (nL - nR) / h and dF/dv = (nL + nR) / 2 (the Bernoulli form's limit)."""
    import mpmath as mp

    D, h, mu, nl, nr, E = 1e3, 1e-3, 1e-3, 5.0, 2.0, 1e-8     # |v| = 1e-11, arg = 1e-17
    for q in (1.0, -1.0):
        assert abs(q * mu * E * h / D) < 2.0 ** -53
        assert O.eval_flux([[E, h, q, mu, D, nl, nr]])[0] == D * (nl - nr) / h
        assert O.eval_dfdph([[E, h, q, mu, D, nl, nr]])[0] == 0.5 * (nl + nr)
    # the limit of the exact derivative (50 digits) at a small argument
    mp.mp.dps = 80
    v = mp.mpf("1e-6")                                          # a = 1e-12

    def flux_v(v):
        a_ = v * mp.mpf(h) / mp.mpf(D)
        return v * (mp.mpf(nl) - mp.mpf(nr) * mp.exp(-a_)) / (1 - mp.exp(-a_))

    dv = mp.mpf("1e-26")
    assert abs(float((flux_v(v + dv) - flux_v(v - dv)) / (2 * dv)) - 0.5 * (nl + nr)) < 1e-10
    # just above the degenerate band the reference's own formula stays (finite)
    E2 = 1e-3                                                   # arg = 1e-15
    f = O.eval_flux([[E2, h, 1.0, mu, D, nl, nr]])[0]
    assert np.isfinite(f) and f != D * (nl - nr) / h


def test_dfdph_is_derivative_of_sg_flux(O):
    import mpmath as mp

    mp.mp.dps = 50
    rng = np.random.default_rng(5)
    for _ in range(200):
        E = float(rng.standard_normal() * 10 ** rng.uniform(-3, 1.5))
        h, D = float(rng.uniform(0.3, 2)), float(10 ** rng.uniform(-1, 2))
        q = float(rng.choice([-1.0, 1.0]))
        mu = float(10 ** rng.uniform(-2, 1))
        n1, n2 = float(10 ** rng.uniform(-2, 4)), float(10 ** rng.uniform(-2, 4))
        a = abs(q * mu * E * h / D)
        if a < 1e-3 or a > 60:
            continue
        got = O.eval_dfdph([[E, h, q, mu, D, n1, n2]])[0]

        def flux_v(v):
            a_ = v * mp.mpf(h) / mp.mpf(D)
            return mp.mpf(D) / mp.mpf(h) * (a_ / (1 - mp.exp(-a_)) * mp.mpf(n1) -
                                            a_ * mp.exp(-a_) / (1 - mp.exp(-a_)) * mp.mpf(n2))

        v0 = mp.mpf(q) * mp.mpf(mu) * mp.mpf(E)
        dv = mp.mpf(10) ** -20 * max(abs(v0), 1)
        ref = (flux_v(v0 + dv) - flux_v(v0 - dv)) / (2 * dv)
        assert abs(got - float(ref)) <= 1e-12 * (abs(float(ref)) + (n1 + n2) / a), (E, h, q, mu, D)


# ---------------------------------------------------------------------------
def test_rk_merson_tableau_on_linear_ode(O):
    """Drive the oracle's stage updates with k_s = lambda * n2 on a uniform field:
    the result must be y0*R(z), R = 1+z+z^2/2+z^3/6+z^4/24+z^5/144 (Merson), and
    the embedded error |dt*(2k1/30-3k3/10+4k4/15-k5/30)| = |z^5/720 * y0|."""
    cfg = oracle_cfg(small_cfg(6, 4))
    o = O.Oracle(cfg)
    lam, dt, y0 = -3.0, 0.05, 2.0
    z = lam * dt
    shape = (o.by + 2, o.P)
    for lev in (1, 2, 3):
        o.set("ni", lev, np.full(shape, y0))
    o.set_scalars({"dt": dt})
    interior = np.zeros(shape, bool)
    interior[1:-1, 1:-1] = True
    nerr = None
    for stage in range(1, 6):
        n2 = o.get("ni", 2)
        k = np.where(interior, lam * n2, 0.0)
        o.set("ki", stage, k)
        e = o.rk_all(stage)
        if stage == 5:
            nerr = e[0]
    F = fractions.Fraction
    zf = F(z).limit_denominator(10 ** 12)
    R = 1 + zf + zf ** 2 / 2 + zf ** 3 / 6 + zf ** 4 / 24 + zf ** 5 / 144
    y1 = o.get("ni", 1)[1:-1, 1:-1]
    assert np.allclose(y1, float(R) * y0, rtol=1e-14, atol=0)
    err = abs(float(zf ** 5 / 720) * y0)
    expected = (o.bx * o.by) * (err / (1e-3 + 1e-3 * y0)) ** 2
    assert abs(nerr - expected) / expected < 1e-9


def test_controller_law_on_recorded_attempts(O):
    cfg = oracle_cfg(small_cfg(16, 12, volts=200.0))
    o = O.Oracle(cfg)
    o.init_gaussian(seed=1)
    o.run(12)
    log = o.log()
    assert len(log) >= 10
    err_prev = 1.0
    for i, rec in enumerate(log):
        dt_stage, err_cur, dt_next, acc = rec[2], rec[3], rec[4], rec[5]
        sc = 0.8 * err_cur ** (-0.175) * err_prev ** 0.1
        sc = min(2.5, max(0.3, sc))
        assert dt_next == pytest.approx(sc * dt_stage, rel=1e-13)
        assert bool(acc) == (err_cur <= 1.0)
        if acc:
            err_prev = err_cur
        if i + 1 < len(log) and not acc:  # internal restart reuses the new dt
            assert log[i + 1][2] == dt_next
    assert any(not r[5] for r in log), "expected internal restarts from dt0"


# ---------------------------------------------------------------------------
def test_uniform_and_geometric_grids(O):
    c = oracle_cfg(small_cfg(10, 7, grid_x=0, ratio_x=1.0))
    g = O.Oracle(c).grid()
    i = np.arange(12)
    assert np.allclose(g["x"], c["l"] * i / 11, rtol=1e-15, atol=1e-13)
    assert np.allclose(g["dlx"], 0.5 * (g["dx"][1:] + g["dx"][:-1]))
    c2 = oracle_cfg(small_cfg(10, 7, grid_x=2, ratio_x=1.05))
    g2 = O.Oracle(c2).grid()
    dx = g2["dx"]
    assert dx.size == 11 and abs(dx.sum() - c2["l"]) < 1e-12 * c2["l"]
    assert np.allclose(dx[1:6] / dx[:5], 1.05, rtol=1e-13)
    assert np.allclose(dx, dx[::-1], rtol=1e-13)
    # node types (props.F90:186-218): full plates -> electrode at both x ends
    tx, ty = g2["type_x"], g2["type_y"]
    assert (tx[:, 0] == -2).all() and (tx[:, -1] == 2).all() and (tx[:, 1:-1] == 0).all()
    assert (ty[0] == -1).all() and (ty[-1] == 1).all() and (ty[1:-1] == 0).all()


def test_partial_electrode_types(O):
    c = oracle_cfg(small_cfg(8, 10))
    c["ew"] = c["w"] * 0.3
    c["gap"] = c["w"] * 0.75
    g = O.Oracle(c).grid()
    y = g["y"]
    for j in range(10):
        el = (y[j] <= c["ew"]) or (y[j] >= c["gap"])  # y(j) for cell j+1 (props.F90:201)
        assert g["type_x"][j, 0] == (-2 if el else -1)
        assert g["type_x"][j, -1] == (2 if el else 1)


# ---------------------------------------------------------------------------
def test_rhs_uniform_state_sources_and_electrode_flux(O):
    """phi = 0 and uniform densities: every interior face flux vanishes, so the
    slope is the source term alone and equal in all cells except those at the
    electrodes, which receive the thermal wall fluxes (ptcl_lib.F90:201-212)."""
    c = oracle_cfg(small_cfg(9, 6))
    o = O.Oracle(c)
    k = O.consts(c["p_torr"], c["gam"])
    shape = (o.by + 2, o.P)
    ni, ne, nm = 3.0, 2.0, 1.5
    Te = 200.0
    for lev in (1, 2, 3):
        o.set("ni", lev, np.full(shape, ni))
        o.set("ne", lev, np.full(shape, ne))
        o.set("nm", lev, np.full(shape, nm))
        o.set("nt", lev, np.full(shape, ne * Te))
    o.set("ph", 1, np.zeros(shape))
    o.rhs_sweep(1)
    ki = o.get("ki", 1)[1:-1, 1:-1]
    km = o.get("km", 1)[1:-1, 1:-1]
    g = o.grid()
    inner = ki[:, 1:-1]
    assert np.all(inner == inner[0, 0])
    vi = k["vi"]
    # left cell: k = (F_R - F_L)*(-1)/dlx + S with F_R = 0, F_L = -vi/4*ni (a = 0)
    left = ki[:, 0] - inner[0, 0]
    assert np.allclose(left, -(0.25 * vi * ni) / g["dlx"][0], rtol=1e-12)
    right = ki[:, -1] - inner[0, 0]
    assert np.allclose(right, -(0.25 * vi * ni) / g["dlx"][-1], rtol=1e-12)
    # metastables: wall loss vi/2*nm on both electrodes
    assert np.allclose(km[:, 0] - km[:, 3], -(0.5 * vi * nm) / g["dlx"][0], rtol=1e-12)


def test_rhs_is_y_invariant_for_x_only_state(O):
    c = oracle_cfg(small_cfg(14, 9))
    o = O.Oracle(c)
    g = o.grid()
    x = g["x"]
    shape = (o.by + 2, o.P)
    prof = lambda a, b: np.tile(a + b * np.sin(np.pi * x / x[-1]), (shape[0], 1))
    for lev in (1, 2, 3):
        o.set("ni", lev, prof(5.0, 40.0))
        o.set("ne", lev, prof(4.0, 38.0))
        o.set("nm", lev, prof(1.0, 3.0))
        o.set("nt", lev, prof(4.0, 38.0) * 150.0)
    o.set("ph", 1, prof(0.0, 3e3))
    o.rhs_sweep(2)
    for name in ("ki", "ke", "kt", "km"):
        k = o.get(name, 2)[1:-1, 1:-1]
        assert np.allclose(k, k[0][None, :], rtol=1e-12, atol=0), name


def test_rhs_interior_faces_telescope(O):
    """sum_cells V*(ki - ke) = -sum over the two electrode columns of dly*(flux)
    because every interior face enters two cells with opposite signs."""
    c = oracle_cfg(small_cfg(12, 8))
    o = O.Oracle(c)
    o.init_gaussian(seed=4, noise=0.3)
    o.run(3)
    g = o.grid()
    o.rhs_sweep(1)
    ki = o.get("ki", 1)[1:-1, 1:-1]
    ke = o.get("ke", 1)[1:-1, 1:-1]
    V = np.outer(g["dly"], g["dlx"])
    total = np.sum(V * (ki - ke))
    # electrode face fluxes from the same state (closed-form wall BC)
    k = O.consts(c["p_torr"], c["gam"])
    n2 = {nm: o.get(nm, 2) for nm in ("ni", "ne", "nt")}
    ph = o.get("ph")
    fe = O.eval_fits(np.array([1.0]), c["p_torr"])  # noqa: F841 (exercise)
    wall = 0.0
    for side in ("L", "R"):
        ic = 1 if side == "L" else o.bx
        ig = 0 if side == "L" else o.bx + 1
        h = g["dx"][0] if side == "L" else g["dx"][-1]
        for j in range(1, o.by + 1):
            E = -(ph[j, ic] - ph[j, ig]) / h if side == "L" else -(ph[j, ig] - ph[j, ic]) / h
            ni, ne, nt = n2["ni"][j, ic], n2["ne"][j, ic], n2["nt"][j, ic]
            Te = nt / ne
            a = 1.0 if ph[j, ig] > ph[j, ic] else 0.0
            mue = O.eval_fits(np.array([Te]), c["p_torr"])[0, 0]
            ve = np.sqrt((16.0 * k["e"] * k["ph0"] * Te) / (3.0 * k["pi"] * k["me"])) * k["t0"] / k["x0"]
            sgn = -1.0 if side == "L" else 1.0
            Fi = (1.0 - a) * k["mui"] * E * ni + sgn * 0.25 * k["vi"] * ni
            Fe = -(a * mue * E * ne) + sgn * 0.25 * ve * ne - k["gam"] * Fi
            # outward flux leaves the domain: contributes -(F_R - F_L)
            wall += (-1.0 if side == "R" else 1.0) * g["dly"][j - 1] * (Fi - Fe)
    scale = np.sum(np.abs(V * ki)) + np.sum(np.abs(V * ke))
    assert abs(total - wall) <= 1e-11 * scale, (total, wall)


# ---------------------------------------------------------------------------
def _assemble(o):
    st, b, A = o.lapl_assemble()
    assert st == 0
    return b, A


def _apply(A, phi_int):
    ny, nx = phi_int.shape
    out = A[..., 4] * phi_int
    out[:, 1:] += A[:, 1:, 0] * phi_int[:, :-1]
    out[:, :-1] += A[:, :-1, 1] * phi_int[:, 1:]
    out[1:, :] += A[1:, :, 2] * phi_int[:-1, :]
    out[:-1, :] += A[:-1, :, 3] * phi_int[1:, :]
    return out


def test_laplJ_is_jacobian_of_laplEqn(O):
    c = oracle_cfg(small_cfg(10, 7), solver=0)
    o = O.Oracle(c)
    o.init_gaussian(seed=2, noise=0.2)
    o.run(4)
    b0, A = _assemble(o)
    ph = o.get("ph")
    rng = np.random.default_rng(0)
    for _ in range(12):
        j, i = int(rng.integers(1, o.by + 1)), int(rng.integers(1, o.bx + 1))
        eps = 1e-5 * max(1.0, abs(ph[j, i]))
        cols = []
        for sgn in (1, -1):
            p2 = ph.copy()
            p2[j, i] += sgn * eps
            o.set("ph", 1, p2)
            cols.append(_assemble(o)[0])
        o.set("ph", 1, ph)
        fd = (cols[0] - cols[1]) / (2 * eps)
        e = np.zeros((o.by, o.bx))
        e[j - 1, i - 1] = 1.0
        an = _apply(A, e)
        scale = np.abs(an).max()
        assert np.abs(fd - an).max() <= 1e-5 * scale, (j, i)


def test_laplacian_when_dt_is_zero(O):
    c = oracle_cfg(small_cfg(9, 6, grid_x=2, ratio_x=1.05))
    o = O.Oracle(c)
    o.init_gaussian(seed=1)
    o.set_scalars({"dt": 0.0})
    g = o.grid()
    ph = o.get("ph")
    ph[1:-1, 1:-1] = np.random.default_rng(1).standard_normal((o.by, o.bx)) * 10
    o.set("ph", 1, ph)
    b, A = _assemble(o)
    dx, dlx, dy, dly = g["dx"], g["dlx"], g["dy"], g["dly"]
    for jj in range(o.by):
        for ii in range(o.bx):
            if 0 < ii < o.bx - 1:
                assert A[jj, ii, 0] == pytest.approx(1 / dx[ii] / dlx[ii], rel=1e-15)
                assert A[jj, ii, 1] == pytest.approx(1 / dx[ii + 1] / dlx[ii], rel=1e-15)
            if 0 < jj < o.by - 1:
                assert A[jj, ii, 2] == pytest.approx(1 / dy[jj] / dly[jj], rel=1e-15)
                assert A[jj, ii, 4] == pytest.approx(
                    -(1 / dx[ii + 1] + 1 / dx[ii]) / dlx[ii] - (1 / dy[jj + 1] + 1 / dy[jj]) / dly[jj],
                    rel=1e-14)
    ne, ni = o.get("ne", 1), o.get("ni", 1)
    lap = _apply(A, ph[1:-1, 1:-1])
    # Dirichlet ghosts at the electrodes (zero potential) enter through Ex(1)
    assert np.allclose(b, lap + ni[1:-1, 1:-1] - ne[1:-1, 1:-1], rtol=1e-12, atol=1e-9)


def test_volume_weighted_operator_is_symmetric(O):
    c = oracle_cfg(small_cfg(11, 8))
    o = O.Oracle(c)
    o.init_gaussian(seed=3, noise=0.2)
    o.run(4)
    _, A = _assemble(o)
    g = o.grid()
    V = np.outer(g["dly"], g["dlx"])
    W = V[:, 1:] * A[:, 1:, 0]
    E = V[:, :-1] * A[:, :-1, 1]
    assert np.allclose(W, E, rtol=1e-13)
    S = V[1:, :] * A[1:, :, 2]
    N = V[:-1, :] * A[:-1, :, 3]
    assert np.allclose(S, N, rtol=1e-13)
    assert (A[..., :4] >= 0).all() and (A[..., 4] < 0).all()


def test_poisson_solvers_agree(O):
    base = small_cfg(14, 9)
    sols = {}
    for solver in (0, 1, 2):
        o = O.Oracle(oracle_cfg(base, solver=solver, rtol=1e-13))
        o.init_gaussian(seed=6)
        o.run(3)
        sols[solver] = o.get("ph")
        b, A = _assemble(o)
    # banded LU against a dense solve of the same system
    o = O.Oracle(oracle_cfg(base, solver=0))
    o.init_gaussian(seed=6)  # first solve: phi = 0, so delta is the whole potential
    b, A = _assemble(o)
    N = o.bx * o.by
    M = np.zeros((N, N))
    for n in range(N):
        e = np.zeros(N)
        e[n] = 1
        M[:, n] = _apply(A, e.reshape(o.by, o.bx)).ravel()
    x = np.linalg.solve(M, -b.ravel())
    ph0 = o.get("ph")[1:-1, 1:-1].copy()
    st, lerr, stats = o.lapl_solve()
    assert lerr == 0
    assert np.allclose(o.get("ph")[1:-1, 1:-1] - ph0, x.reshape(o.by, o.bx), rtol=1e-9,
                       atol=1e-9 * np.abs(x).max())
    for s in (1, 2):
        assert np.allclose(sols[s], sols[0], rtol=1e-9, atol=1e-9 * np.abs(sols[0]).max())


# ---------------------------------------------------------------------------
def test_reject_path_restores_levels_and_halves_dt(O):
    c = oracle_cfg(small_cfg(10, 8), solver=2, rtol=1e-30)
    c["maxit"] = 2  # the solve cannot converge -> lerr = 1 (lapl_lib.F90:106-108)
    o = O.Oracle(c)
    o.init_gaussian(seed=1)
    s0 = o.scalars()
    ph0 = o.get("ph")
    n3 = o.get("ne", 3)
    o.run(1)
    s1 = o.scalars()
    assert s1["dt"] == s0["dt"] / 2 and s1["ts"] == 0 and s1["t"] == s0["t"]
    assert np.array_equal(o.get("ph"), ph0)
    assert np.array_equal(o.get("ne", 1), n3) and np.array_equal(o.get("ne", 2), n3)


def test_fatal_paths(O):
    c = oracle_cfg(small_cfg(8, 6))
    o = O.Oracle(c)
    o.init_gaussian(seed=1)
    o.run(2)
    nt = o.get("nt", 2)
    nt[3, 4] = -1.0  # negative energy density -> get_Te stops
    o.set("nt", 2, nt)
    assert o.ptcl_step() == -2
    o = O.Oracle(c)
    o.init_gaussian(seed=1)
    o.run(2)
    ne = o.get("ne", 2)
    ne[2, 2] = np.nan
    o.set("ne", 2, ne)
    assert o.ptcl_step() in (-2, -3)
    o = O.Oracle(c)
    o.init_gaussian(seed=1)
    o.run(2)
    o.set_scalars({"dt": 4e-13})  # even 2.5x growth keeps dt below 1.1e-12
    assert o.ptcl_step() == -4
