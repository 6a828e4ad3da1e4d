"""The stage kernel's short arithmetic (fsim_phys.h), compiled for the host from
the same header: fs_log and fs_exp against long-double libm, fs_div / fs_rcp
against long-double division, and the Horner-form fit exponents against the
reference's power-sum order."""
import os
import re
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.fixture(scope="module")
def report(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    exe = str(tmp_path_factory.mktemp("stage_math") / "check")
    src = os.path.join(HERE, "csrc", "stage_math_check.hip")
    # the product's flags: no contraction
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-O2", "-std=c++17", "-ffp-contract=off",
                    "-I" + os.path.join(ROOT, "include"),
                    "-I" + os.path.join(ROOT, "fluidsim_amd", "csrc"), src, "-o", exe],
                   check=True, capture_output=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    rows = {}
    for line in out.splitlines():
        m = re.match(r"(\w+) \[(\S+), (\S+)\] max_ulp (\S+) at \S+ not_nearest (\S+)", line)
        rows[(m.group(1), float(m.group(2)), float(m.group(3)))] = (float(m.group(4)),
                                                                     float(m.group(5)))
    return rows


def _rows(report, fn):
    return {k: v for k, v in report.items() if k[0] == fn}


def test_fs_log_within_one_ulp(report):
    rows = _rows(report, "log")
    assert len(rows) == 3
    for key, (ulp, wrong) in rows.items():
        # within one ulp everywhere, correctly rounded for nearly all arguments
        assert ulp < 1.0, (key, ulp)
        assert wrong < 0.08, (key, wrong)


def test_fs_exp_within_one_ulp(report):
    rows = _rows(report, "exp")
    assert len(rows) == 3
    for key, (ulp, wrong) in rows.items():
        assert ulp < 1.0, (key, ulp)
        assert wrong < 0.1, (key, wrong)


def test_newton_division_and_reciprocal(report):
    for fn in ("div", "rcp"):
        (key, (ulp, wrong)), = _rows(report, fn).items()
        assert ulp <= 0.5 + 1e-9 and wrong < 0.01, (key, ulp, wrong)


# largest |Horner - power sum| of each fit's exponent, in units of 2^-52 (the
# relative difference of the fit itself): the terms of the power sums reach
# 8e3 against values of 20 .. 170, so the two orders round differently by up
# to 6e-12; the device fit parity test bounds the stage fits at 1e-11
FIT_BOUND = {"fit_mue": 2048, "fit_De": 32768, "fit_kex": 8192, "fit_kir": 8192,
             "fit_nu": 2048, "fit_ksc": 1024, "fit_ksi": 1024}


def test_horner_fits_against_the_reference_order(report):
    for name, bound in FIT_BOUND.items():
        (key, (diff, _)), = _rows(report, name).items()
        assert diff <= bound, (key, diff)
