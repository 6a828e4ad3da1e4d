"""The Fortran host (fluidsim_amd/fortran): the reference's module interfaces
(props, ptcl_lib, lapl_lib, circ_lib, main) over libfsim_hip through
ISO_C_BINDING.  It writes the reference's raw FP64 output files."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import oracle_cfg, rel_l2

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "fluidsim_amd", "fortran", "_build", "fsim_main")
HAVE_FC = shutil.which("amdflang") or os.path.exists("/opt/rocm/bin/amdflang")


@pytest.mark.skipif(not HAVE_FC, reason="amdflang not available")
def test_fortran_host_builds_and_links():
    from fluidsim_amd import build

    exe = build.build_fortran()
    assert exe and os.path.exists(exe)
    out = subprocess.run(["nm", "-u", exe], capture_output=True, text=True).stdout
    for sym in ("fs_create", "fs_ptcl_step", "fs_lapl_solve", "fs_accept", "fs_reject",
                "fs_circ_init", "fs_set_field", "fs_get_field", "fs_set_scalars"):
        assert sym in out


@pytest.mark.skipif(not HAVE_FC, reason="amdflang not available")
def test_fortran_driver_flag_errors():
    """read_in's checks stop before any device call: an odd argument count
    prints the usage, px * py /= 1, ny < 1 and unknown flags stop with their
    codes."""
    from fluidsim_amd import build

    exe = build.build_fortran()
    r = subprocess.run([exe, "-nx"], capture_output=True, text=True, timeout=60)
    assert "Usage" in r.stdout and "-wall" in r.stdout and "-n0" in r.stdout
    for args, code in ((["-px", "2"], 2), (["-ny", "0"], 3), (["-bogus", "1"], 2)):
        r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=60)
        assert r.returncode == code, (args, r.returncode, r.stdout)


SERIES = ("time.dat", "dt.dat", "vd.dat", "ida.dat", "idc.dat", "flx.dat", "meshx.dat",
          "meshy.dat")


def _read_frames(path, name, n):
    a = np.fromfile(os.path.join(path, name), dtype=np.float64)
    return a.reshape(-1, n)


def _run(args, out):
    cmd = [EXE, *args, "-out", out]
    print("running", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=180)
    print(r.stdout[-1500:], r.stderr[-1500:])
    assert r.returncode == 0
    return r.stdout


def _same_outputs(a, b, n):
    for f in ("f1.dat", "f2.dat", "f3.dat", "f4.dat", "f5.dat"):
        fa, fb = _read_frames(a, f, n), _read_frames(b, f, n)
        assert fa.shape == fb.shape and np.array_equal(fa, fb), f
    for f in SERIES:
        assert np.array_equal(np.fromfile(a + f), np.fromfile(b + f)), f


@pytest.mark.gpu
def test_fortran_driver_matches_library_and_oracle(O, tmp_path):
    """The reference's call sequence (g_init, lapl_init, ptcl_init,
    circ_init, sfc_init; ptcl_step(g, ph, Em), lapl_solve(g, ne(:,:,1),
    ni(:,:,1), nt(:,:,1), sig_pl, lerr); the accept / reject block on the
    module arrays) against the library's own loop, bitwise, and the oracle;
    the device-resident mode writes the same files bit for bit."""
    if not os.path.exists(EXE):
        pytest.skip("Fortran host not built")
    from fluidsim_amd import Discharge, configs

    nx, ny, steps, seed = 24, 16, 8, 3
    out = str(tmp_path) + "/mirror/"
    os.makedirs(out)
    args = ["-nx", str(nx), "-ny", str(ny), "-p", "1.0", "-v", "200", "-r", "1e5",
            "-grid", "2", "-ratio", "1.05", "-ic", "1", "-seed", str(seed), "-steps",
            str(steps), "-rtol", "1e-13", "-l", "0.01", "-w", "0.01", "-ew", "0.01"]
    stdout = _run(args + ["-verbose", "T"], out)
    for key in ("Timestep:", "time/us:", "tm:", "Em:", "Vd:", "Ida:", "Idc:", "wp:", "ne:"):
        assert key in stdout, key
    f2 = _read_frames(out, "f2.dat", nx * ny)[-1].reshape(ny, nx)   # ne [m^-3]
    f1 = _read_frames(out, "f1.dat", nx * ny)[-1].reshape(ny, nx)   # phi [V]
    t = np.fromfile(out + "time.dat")
    meshx = np.fromfile(out + "meshx.dat")

    cfg = configs.base(nx, ny, p_torr=1.0, volts=200.0, res_ohm=1e5, rtol=1e-13)
    d = Discharge(cfg)
    d.init_gaussian(seed=seed)
    while d.scalars()["ts"] < steps:
        d.step()
    k = d.scalars()
    x0 = configs.X0
    ne = d.get("ne", 1)[1:-1, 1:-1] / x0 ** 3
    ph = d.get("ph")[1:-1, 1:-1] * configs.PH0
    assert np.array_equal(f2, ne) and np.array_equal(f1, ph)
    assert t[-1] == k["t"]
    g = d.grid()
    assert np.array_equal(meshx, g["x"][1:-1] * x0)

    res = str(tmp_path) + "/resident/"
    os.makedirs(res)
    _run(args + ["-resident", "T"], res)
    _same_outputs(out, res, nx * ny)

    o = O.Oracle(oracle_cfg(cfg))
    o.init_gaussian(seed=seed)
    while o.scalars()["ts"] < steps:
        o.run(1)
    assert rel_l2(f2, o.get("ne", 1)[1:-1, 1:-1] / x0 ** 3) < 1e-9


@pytest.mark.gpu
def test_fortran_uniform_start_and_dielectric_wall(tmp_path):
    """ptcl_init's uniform state with -n0 (ptcl_lib.F90:22-29, main.F90:
    240-243) and the dielectric wall (-wall T): the host's accept / reject
    block, including its rotation of sig_pl / sig_mi / sig_or, against the
    library's device-side accept / reject, bitwise, over a run with
    rejected attempts; and the device-resident mode."""
    if not os.path.exists(EXE):
        pytest.skip("Fortran host not built")
    from fluidsim_amd import Discharge, configs
    from fluidsim_amd.discharge import constants

    nx, ny, steps, n0 = 24, 20, 30, 3e10
    args = ["-nx", str(nx), "-ny", str(ny), "-p", "1.0", "-v", "300", "-r", "1e5",
            "-ic", "0", "-n0", str(n0), "-wall", "T", "-steps", str(steps), "-rtol", "1e-13",
            "-l", "0.01", "-w", "0.01", "-ew", "0.01"]
    out = str(tmp_path) + "/mirror/"
    os.makedirs(out)
    _run(args, out)

    cfg = configs.base(nx, ny, p_torr=1.0, volts=300.0, res_ohm=1e5, rtol=1e-13, rwall=1,
                       grid_x=0)
    d = Discharge(cfg)
    tab = constants(1.0, 0.1)
    n = n0 * configs.X0 ** 3
    Tg, kb, e, ph0 = tab[11], tab[4], tab[3], tab[7]
    shape = (ny + 2, nx + 2)
    for lev in (1, 2, 3):
        for name in ("ni", "ne", "nm"):
            d.set(name, lev, np.full(shape, n))
        d.set("nt", lev, np.full(shape, n) * Tg * kb / e / ph0)
    while d.scalars()["ts"] < steps:
        d.step()
    log = d.log()
    print("attempts", len(log), "rejected", int(np.sum(log[:, 5] == 0)))
    assert np.any(log[:, 5] == 0)
    x0 = configs.X0
    f2 = _read_frames(out, "f2.dat", nx * ny)[-1].reshape(ny, nx)
    f1 = _read_frames(out, "f1.dat", nx * ny)[-1].reshape(ny, nx)
    assert np.array_equal(f2, d.get("ne", 1)[1:-1, 1:-1] / x0 ** 3)
    assert np.array_equal(f1, d.get("ph")[1:-1, 1:-1] * configs.PH0)
    assert np.abs(d.get_sig("sig_pl")).max() > 0
    assert np.fromfile(out + "time.dat")[-1] == d.scalars()["t"]

    res = str(tmp_path) + "/resident/"
    os.makedirs(res)
    _run(args + ["-resident", "T"], res)
    _same_outputs(out, res, nx * ny)


@pytest.mark.gpu
def test_fortran_one_dimensional_default(tmp_path):
    """The reference's default mode, ny = 1 (main.F90:29): no -ny flag, the
    uniform start, the DC source; the host's accept / reject block on the
    module arrays against the library's loop bitwise, the device-resident
    mode, and meshy.dat left empty (props.F90:262-264)."""
    if not os.path.exists(EXE):
        pytest.skip("Fortran host not built")
    from fluidsim_amd import Discharge, configs
    from fluidsim_amd.discharge import constants

    nx, steps, n0 = 40, 40, 3e10
    args = ["-nx", str(nx), "-p", "1.0", "-v", "200", "-r", "1e5", "-ic", "0", "-n0", str(n0),
            "-steps", str(steps), "-rtol", "1e-13", "-l", "0.01", "-w", "0.01", "-ew", "0.01"]
    out = str(tmp_path) + "/mirror/"
    os.makedirs(out)
    _run(args, out)
    assert os.path.getsize(out + "meshy.dat") == 0

    cfg = configs.base(nx, 1, p_torr=1.0, volts=200.0, res_ohm=1e5, rtol=1e-13, grid_x=0)
    d = Discharge(cfg)
    tab = constants(1.0, 0.1)
    n = n0 * configs.X0 ** 3
    Tg, kb, e, ph0 = tab[11], tab[4], tab[3], tab[7]
    for lev in (1, 2, 3):
        for name in ("ni", "ne", "nm"):
            d.set(name, lev, np.full((3, nx + 2), n))
        d.set("nt", lev, np.full((3, nx + 2), n) * Tg * kb / e / ph0)
    while d.scalars()["ts"] < steps:
        d.step()
    x0 = configs.X0
    f2 = _read_frames(out, "f2.dat", nx)[-1]
    f1 = _read_frames(out, "f1.dat", nx)[-1]
    assert np.array_equal(f2, d.get("ne", 1)[1, 1:-1] / x0 ** 3)
    assert np.array_equal(f1, d.get("ph")[1, 1:-1] * configs.PH0)
    assert np.fromfile(out + "time.dat")[-1] == d.scalars()["t"]
    assert np.fromfile(out + "ida.dat")[-1] != 0.0

    res = str(tmp_path) + "/resident/"
    os.makedirs(res)
    _run(args + ["-resident", "T"], res)
    _same_outputs(out, res, nx)
