"""The stage kernel's logarithm (fs_log in fsim_phys.h), compiled for the host
from the same header and compared with long-double libm."""
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


def test_fs_log_within_one_ulp(report):
    assert len(report) == 3
    for key, (ulp, wrong) in report.items():
        # within one ulp everywhere, correctly rounded for nearly all arguments
        assert ulp < 1.0, (key, ulp)
        assert wrong < 0.08, (key, wrong)
