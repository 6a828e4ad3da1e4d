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
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|const char\*)\s+(fs_\w+)\s*\(", text, re.M)))


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
