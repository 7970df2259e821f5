"""The committed oracle fixtures under tests/golden/ are what the oracle
produces now (CPU only; the short fixtures are re-run bit for bit, the long
Couette run is checked for its shape and symmetry)."""
import os
import sys

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLDEN)
import make_golden  # noqa: E402

pytest.importorskip("pyoracle")


def _load(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)


@pytest.mark.parametrize("name", ["plate128x64_rk3_nnd", "shear48_rk1_weno5", "riemann96_rk3_weno5"])
def test_short_fixture_reproduces(name):
    z = _load(name)
    case = make_golden.build_case(str(z["spec"]))
    st, m = make_golden.run_oracle(case, int(z["steps"]))
    assert st == 0
    for k in make_golden.FIELDS:
        np.testing.assert_array_equal(m[k], z[k], err_msg=k)


def test_couette_fixture_is_x_invariant_and_sheared():
    z = _load("couette64_rk3_weno5")
    assert int(z["steps"]) == 2000
    ux = z["ux"]
    assert ux.shape == (64, 64)
    # periodic in x with x-invariant walls: every column is the same
    assert np.all(ux == ux[:1, :])
    # wall-normal velocity: small (the compressible transient), not zero
    assert np.max(np.abs(z["uy"])) < 0.01 * 0.1
    # fluid velocity rises from the resting wall (y = 0) to the moving one
    prof = ux[0, 1:-1]
    assert np.all(np.diff(prof) > -0.01) and prof[-1] - prof[0] > 0.05
    assert 0 < float(z["floor"]) < 1e-11


def test_failure_fixture_is_recorded():
    z = _load("couette64_uq84_tc1_failure")
    assert 0 < int(z["fail_step"]) < int(z["max_steps"])
    case = make_golden.build_case(str(z["spec"]))
    assert case.params["Tc_Tref"] == 1.0 and case.meta["mach"] > 0.4


def test_jacobs_digitisation_meets_the_deck():
    """The digitised Jacobs profiles, scaled as the reference scales them,
    end at the reference deck's free stream: u -> 597.3 m/s and T -> 222 K
    (a wrong axis scale or flipped row direction would miss both)."""
    import json
    from uq_amd import deck
    d = json.load(open(os.path.join(GOLDEN, "jacobs_profiles.json")))
    c = deck.read_mat_deck(os.path.join(GOLDEN, "jacobs_boundary_layer_input.mat"))
    u_inf, t_inf = c.props["velocity_x"][0], c.props["temp"][0]
    u = np.array(d["ux"]["value_m_s"])
    t = np.array(d["T"]["value_K"])
    assert abs(u[-5:].mean() - u_inf) < 0.01 * u_inf
    assert abs(t[-4:].mean() - t_inf) < 0.5
    assert u[0] < 5 and np.all(np.diff(u[:25]) > 0)          # no-slip wall, rising profile
    assert t.max() > t_inf + 30                               # heating inside the layer
