"""Deck and snapshot formats (CPU only): the reference's MAT deck reads
into the solver's deck with loadInput's variable mapping and y-flip, and a
snapshot round-trips through the createMAT layout."""
import os

import numpy as np
import pytest

from uq_amd import cases, deck

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
JACOBS = os.path.join(GOLDEN, "jacobs_boundary_layer_input.mat")


def test_reads_reference_boundary_layer_deck():
    c = deck.read_mat_deck(JACOBS)
    p = c.params
    assert (c.nx, c.ny) == (1001, 101)
    assert p["rk_method"] == 1 and p["flux_method"] == 1
    assert p["cfl"] == pytest.approx(0.3) and p["tt_time"] == pytest.approx(0.01)
    assert p["R"] == 287 and p["gamma"] == pytest.approx(1.4) and p["Pr"] == pytest.approx(0.72)
    assert p["dx"] == pytest.approx(1e-3) and p["dy"] == pytest.approx(1e-3)
    assert p["T_ref"] == 293 and p["Tc_Tref"] == pytest.approx(2.5)
    assert p["mu"] == pytest.approx(1.46254376e-05)
    assert list(c.props["cell_type"]) == [1, 0, 2]
    assert c.props["velocity_x"][0] == pytest.approx(597.32604162)
    # after the flip the wall (id 2) is row y = 0, the permanent free stream
    # (id 0) is the top row and the inflow column
    assert np.all(c.bnd[:, 0] == 2)
    assert np.all(c.bnd[:, c.ny - 1] == 0)
    assert np.all(c.bnd[0, 1:] == 0)
    assert np.all(c.bnd[1:, 1:c.ny - 1] == 1)


def test_deck_runs_on_the_oracle():
    pyoracle = pytest.importorskip("pyoracle")
    c = deck.read_mat_deck(JACOBS)
    # a 40 x 12 corner of the deck keeps its structure (wall row, inflow
    # column, free-stream top row)
    sub = cases.Case("jacobs_corner", dict(c.params), np.ascontiguousarray(c.bnd[:40, -12:]).copy(),
                     c.props, 5)
    sub.bnd[:, 0] = 2
    o = pyoracle.Oracle(sub)
    k = o.consts()
    # dt = CFL*dx/(2c), c = sqrt(R*Tc_Tref*T_ref); nsteps from tt_time
    cc = (287 * 2.5 * 293) ** 0.5
    assert k["dt"] == pytest.approx(0.3 * 1e-3 / (2 * cc))
    assert k["nsteps"] == int(np.ceil(0.01 / k["dt"]))
    assert o.step(3) == 0


def test_snapshot_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    nx, ny = 7, 5
    macro = {k: rng.standard_normal((nx, ny)) for k in ("rho", "ux", "uy", "p", "T")}
    f, h = rng.standard_normal((nx, ny, 13)), rng.standard_normal((nx, ny, 13))
    path = str(tmp_path / "step 10.mat")
    deck.write_mat_snapshot(path, macro, t=0.5, dx=0.1, dy=0.2, dist=(f, h), c=2.0)
    import scipy.io
    raw = scipy.io.loadmat(path)
    assert raw["rho"].shape == (ny, nx)            # {Ny, Nx}
    assert raw["f"].shape == (13, ny, nx)          # {13, Ny, Nx}
    assert raw["rho"][3, 2] == macro["rho"][2, 3]  # element (y, x)
    assert raw["f"][4, 3, 2] == f[2, 3, 4]
    np.testing.assert_array_equal(raw["ex"].ravel(), 2.0 * np.array([0, 1, 0, -1, 0, 1, -1, -1, 1, 2, 0, -2, 0]))
    back = deck.read_mat_snapshot(path)
    for k in macro:
        np.testing.assert_array_equal(back[k], macro[k])
    np.testing.assert_array_equal(back["f"], f)
    assert back["t"] == 0.5 and back["dy"] == 0.2


def test_reference_deck_passes_geometry_validation():
    """The library's host-side check of its stencil addressing accepts the
    full 1001 x 101 reference deck (status OK with a GPU, NOGPU without)."""
    import ctypes as C
    from uq_amd import capi, lbm
    c = deck.read_mat_deck(JACOBS)
    d = capi.PackedDeck(c.params, c.bnd, c.props)
    h = C.c_void_p()
    st = lbm.lib().lbm_create(C.byref(d.params), c.nx, c.ny, d.bnd_ptr(), C.byref(d.props), 0, C.byref(h))
    if st == 0:
        lbm.lib().lbm_destroy(h)
    assert st in (capi.LBM_OK, capi.LBM_ERR_NOGPU), lbm.lib().lbm_last_error().decode()
