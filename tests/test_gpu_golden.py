"""GPU FP64 path against the committed oracle fixtures (tests/golden/),
including the 64 x 64 Couette deck at its full 2000 steps.

Bar: relative L2 <= max(1e-12, 8 x the oracle's own one-ulp sensitivity
over the same run), on rho, T and the velocity vector; the sensitivity is
stored in each fixture by make_golden.py."""
import os
import sys

import numpy as np
import pytest

from uq_amd.lbm import rel_l2

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLDEN)
import make_golden  # noqa: E402


@pytest.mark.parametrize("name", sorted(make_golden.FIXTURES))
def test_gpu_matches_fixture(name, gpu_lib):
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    case = make_golden.build_case(str(z["spec"]))
    steps = int(z["steps"])
    bar = max(1e-12, 8.0 * float(z["floor"]))
    print(f"[golden] {name} {case.nx}x{case.ny} steps={steps} bar={bar:.2e}", flush=True)
    s = gpu_lib.Solver(case, "fp64")
    try:
        assert s.step(steps) == 0
        m = s.get_macro()
    finally:
        s.close()
    ref = {k: z[k] for k in make_golden.FIELDS}
    errs = make_golden.field_errors(m, ref)
    print(f"[golden] {name}: " + " ".join(f"{k}={v:.2e}" for k, v in errs.items()), flush=True)
    for k, v in errs.items():
        assert v <= bar, (k, v, bar)
    if make_golden.FIXTURES[name][4]:
        # x-invariant decks stay x-invariant on the GPU
        assert np.all(m["ux"] == m["ux"][:1, :])


@pytest.mark.parametrize("name", sorted(make_golden.FAILURES))
def test_gpu_fails_at_oracle_step(name, gpu_lib):
    """A deck on which the scheme itself blows up: the GPU FP64 path trips
    the equilibrium check at the same step as the oracle (thousands of
    steps of agreement before it)."""
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    case = make_golden.build_case(str(z["spec"]))
    fail_step = int(z["fail_step"])
    assert fail_step > 0
    print(f"[golden] {name} {case.nx}x{case.ny}: oracle fails at step {fail_step}", flush=True)
    s = gpu_lib.Solver(case, "fp64")
    try:
        st = s.step(int(z["max_steps"]))
        step = s.info()["step"]
    finally:
        s.close()
    print(f"[golden] {name}: GPU status {st} at step {step}", flush=True)
    assert st == 1 and step == fail_step
