"""The deviation-form equilibria that FP32 storage evaluates (eq_dev in
uq_amd/csrc/lbm_device.hpp) against the closed form of the reference's
equilibrium2D (eq_dir; CPP/inlineFunctions.cpp:10-61), on the host: the
FP64 instance of the regrouped polynomials equals closed form minus the
reference equilibrium to FP64 rounding over cells from 1e-6 to 30 % away
from the reference, so a dropped or altered term shows at the size of the
deviation; the FP32 instance errs relative to the deviation. The check
program (tests/eqdev_check.cpp) is built against the same header the
kernels include."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def test_deviation_form_equals_closed_form(tmp_path):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    exe = tmp_path / "eqdev_check"
    build = subprocess.run([HIPCC, "-std=c++17", "-O2", "--offload-arch=gfx950", "-I", os.path.join(ROOT, "include"),
                            os.path.join(ROOT, "tests", "eqdev_check.cpp"), "-o", str(exe)],
                           capture_output=True, text=True, timeout=600)
    assert build.returncode == 0, build.stderr[-2000:]
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    print(run.stdout)
    assert run.returncode == 0, run.stdout
