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
    for sym in ("fs_create", "fs_ptcl_step", "fs_lapl_solve", "fs_accept", "fs_reject"):
        assert sym in out


def _read_frames(path, name, n):
    a = np.fromfile(os.path.join(path, name), dtype=np.float64)
    return a.reshape(-1, n)


@pytest.mark.gpu
def test_fortran_driver_matches_library_and_oracle(O, tmp_path):
    if not os.path.exists(EXE):
        pytest.skip("Fortran host not built")
    from fluidsim_amd import Discharge, configs

    nx, ny, steps, seed = 24, 16, 8, 3
    out = str(tmp_path) + "/"
    args = [EXE, "-nx", str(nx), "-ny", str(ny), "-p", "1.0", "-v", "200", "-r", "1e5",
            "-grid", "2", "-ratio", "1.05", "-ic", "1", "-seed", str(seed), "-steps",
            str(steps), "-rtol", "1e-13", "-l", "0.01", "-w", "0.01", "-ew", "0.01",
            "-out", out]
    print("running", " ".join(args), flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
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

    o = O.Oracle(oracle_cfg(cfg))
    o.init_gaussian(seed=seed)
    while o.scalars()["ts"] < steps:
        o.run(1)
    assert rel_l2(f2, o.get("ne", 1)[1:-1, 1:-1] / x0 ** 3) < 1e-9
