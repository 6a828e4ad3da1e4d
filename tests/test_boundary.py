"""The C-ABI boundary on the CPU: libfsim_hip.so loads, exports exactly what
include/fsim.h declares, and its host-side constant table equals the oracle's.
No device call is made here."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "fsim.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|const char\*|fs_handle\*)\s+(fs_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    import ctypes

    from fluidsim_amd import lib

    L = lib.load()
    for name in header_functions():
        assert hasattr(L, name), name
        assert isinstance(getattr(L, name), ctypes._CFuncPtr)
    assert sorted(lib.EXPORTS) == header_functions()


def test_exported_symbols_are_exactly_the_header():
    import subprocess

    from fluidsim_amd import lib

    out = subprocess.run(["nm", "-D", "--defined-only", lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = sorted({ln.split()[-1] for ln in out.splitlines()
                       if " T " in ln and ln.split()[-1].startswith("fs_")})
    assert exported == header_functions()


@pytest.mark.parametrize("p,gam", [(10.0, 0.1), (2.0, 0.1), (0.5, 0.05), (7.3, 0.15)])
def test_constant_table_matches_oracle_bitwise(O, p, gam):
    from fluidsim_amd import constants

    a = constants(p, gam)
    b = O.consts_array(p, gam)
    assert np.array_equal(a, b), np.nonzero(a != b)


def test_errors_are_reported_not_swallowed():
    from fluidsim_amd import lib

    with pytest.raises(lib.FsError) as e:
        lib.check(-2)
    assert e.value.code == -2
    assert lib.check(1) == 1 and lib.check(0) == 0


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from fluidsim_amd import lib

    monkeypatch.setattr(lib, "_lib", None)
    monkeypatch.setattr(lib, "LIB_PATH", str(tmp_path / "absent.so"))
    with pytest.raises(ImportError):
        lib.load()


def test_workload_nondimensionalisation():
    from fluidsim_amd import configs

    c0 = configs.config0()
    assert c0["nx"] == c0["ny"] == 64 and c0["p_torr"] == 1.0
    assert c0["l"] == pytest.approx(100.0, rel=1e-7)          # 1 cm in x0 units
    assert c0["vset"] * configs.PH0 == pytest.approx(200.0)    # volts
    c1 = configs.config1()
    assert c1["rf"] == 4 and c1["t_rise"] == 0.01 and c1["t_width"] == 0.1
    assert c1["ratio_x"] ** (512 / 2) == pytest.approx(1.05 ** 32)
    c2 = configs.config2()
    assert c2["l"] == pytest.approx(500.0, rel=1e-7) and c2["grid_x"] == 0


def test_config3_draws_follow_the_sweep_ranges():
    import numpy as np

    from fluidsim_amd import configs
    from fluidsim_amd.configs import PH0

    cfgs = configs.config3(256)
    nx = np.array([c["nx"] for c in cfgs])
    ny = np.array([c["ny"] for c in cfgs])
    v = np.array([c["vset"] * PH0 for c in cfgs])
    p = np.array([c["p_torr"] for c in cfgs])
    g = np.array([c["gam"] for c in cfgs])
    assert nx.min() >= 64 and nx.max() <= 256 and ny.min() >= 64 and ny.max() <= 256
    assert (nx != ny).mean() > 0.9
    assert v.min() >= 150 and v.max() <= 400
    assert p.min() >= 0.5 and p.max() <= 10 and np.median(np.log(p)) == pytest.approx(np.log(np.sqrt(5)), abs=0.3)
    assert g.min() >= 0.05 and g.max() <= 0.15
    assert [c["seed"] for c in cfgs] == list(range(256))
