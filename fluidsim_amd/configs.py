"""Named discharge workloads (physical units -> nondimensional inputs).

Nondimensionalisation follows props.F90:33-37: lengths in x0 = 1e-4 m, times
in t0 = 1 us, potentials in ph0 = e/(eps0*x0) volts; the float32-widened x0,
e and eps0 of the reference are used (main.F90:33-43 divides by x0 the same way).
"""
from __future__ import annotations

import numpy as np

X0 = float(np.float32(1e-4))
E = float(np.float32(1.60217646e-19))
EPS0 = float(np.float32(8.85418782e-12))
PH0 = E / (EPS0 * X0)
X0_3 = X0 * X0 * X0
DT0 = float(np.float32(5.0e-6))  # main.F90:37


def base(nx, ny, *, l_m=1e-2, w_m=1e-2, p_torr=1.0, volts=200.0, res_ohm=1e5, rf=0,
         circuit=0, grid_x=2, ratio_x=1.05, grid_y=0, gam=0.1, cyl=0, rtol=1e-10,
         maxit=20000, t_rise=0.0, t_width=0.0, t_fall=0.0, chunk=32, debug=0,
         solver=1, graphs=1, warm=0, rwall=0, stage_tile=0) -> dict:
    """A parallel-plate argon discharge: electrodes at x = 0 and x = l over the
    full width w (ew = w), a Neumann wall at y = 0 and at y = w either a
    Neumann wall or, with rwall = 1, a dielectric wall that collects surface
    charge (props.F90:47, 192-198)."""
    l = l_m / X0
    w = w_m / X0
    return dict(
        nx=nx, ny=ny, grid_x=grid_x, grid_y=grid_y, ratio_x=ratio_x, ratio_y=1.0,
        l=l, w=w, ew=w, gap=1e-3 / X0, cyl=cyl, rwall=rwall, rf=rf, p_torr=p_torr, gam=gam,
        vset=volts / PH0, res_ohm=res_ohm, t_rise=t_rise, t_width=t_width, t_fall=t_fall,
        circuit=circuit, dt0=DT0, lapl_rtol=rtol, lapl_maxit=maxit, lapl_chunk=chunk,
        debug_keep_k=debug, lapl_solver=solver, use_graphs=graphs, lapl_warm=warm,
        stage_tile=stage_tile,
    )


def config0(**kw) -> dict:
    """64x64, geometric x refinement (ratio 1.05) toward both electrodes, 1 cm gap,
    1 Torr, 200 V DC through a 1e5 ohm ballast, gamma 0.1."""
    return base(64, 64, **kw)


def config1(n=512, **kw) -> dict:
    """512x512, 2 Torr, trapezoid pulse 400 V (10 ns rise, 100 ns flat, 10 ns
    fall).  The geometric ratio is 1.05**(64/n), which keeps the spacing contrast
    of the 64-cell grid (1.05**32); a literal 1.05 per cell at 512 cells would
    make the smallest cell ~1 nm."""
    d = dict(p_torr=2.0, volts=400.0, rf=4, res_ohm=1e5, t_rise=0.01, t_width=0.1,
             t_fall=0.01, ratio_x=1.05 ** (64.0 / n))
    d.update(kw)
    return base(n, n, **d)


def config2(n=2048, **kw) -> dict:
    """2048x2048, 5 cm x 5 cm uniform, 1 Torr, DC 300 V, ballast 5e4 ohm."""
    d = dict(l_m=5e-2, w_m=5e-2, p_torr=1.0, volts=300.0, res_ohm=5e4, grid_x=0,
             ratio_x=1.0)
    d.update(kw)
    return base(n, n, **d)


def ic_defaults(p_torr=1.0) -> dict:
    """Gaussian blob of 1e8 cm^-3 peak, sigma = 0.1 cm, 1% noise (nondimensional)."""
    return dict(peak=1e14 * X0_3, sigma=1e-3 / X0, noise=0.01)



def config3(n_members=256, base_seed=0, lo=64, hi=256, **kw) -> list:
    """Parameter-sweep ensemble: sides drawn uniformly from lo..hi (independently,
    non-square allowed), DC voltage U[150, 400] V, pressure log-uniform in
    [0.5, 10] Torr, secondary emission U[0.05, 0.15]; member i's initial-noise
    seed is base_seed + i (returned as cfg["seed"])."""
    import numpy as np

    rng = np.random.default_rng(12345 + base_seed)
    out = []
    for i in range(n_members):
        nx, ny = (int(v) for v in rng.integers(lo, hi + 1, size=2))
        volts = float(rng.uniform(150.0, 400.0))
        p = float(np.exp(rng.uniform(np.log(0.5), np.log(10.0))))
        gam = float(rng.uniform(0.05, 0.15))
        c = base(nx, ny, p_torr=p, volts=volts, gam=gam, ratio_x=1.05 ** (64.0 / nx), **kw)
        c["seed"] = base_seed + i
        out.append(c)
    return out
