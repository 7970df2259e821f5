"""Pins of the CPU oracle (CPU only).

The reference ships no tests, golden vectors or runnable build (its C++
needs MATLAB's mat.h/engine.h and Win32 windows.h), so each oracle function
is pinned by what the mathematics fixes: moment identities of the
equilibria, closed forms, invariants, brute force on tiny grids and the
order of accuracy of the time stepper. Each test is chosen so that a
plausible slip (a dropped term, a sign, a transposed index) fails it.
"""
import math
import os

import numpy as np
import pytest

from uq_amd import cases

pyoracle = pytest.importorskip("pyoracle")
Oracle = pyoracle.Oracle

EX = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1, 2, 0, -2, 0], dtype=float)
EY = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1, 0, 2, 0, -2], dtype=float)


@pytest.fixture(scope="module")
def si_oracle():
    return Oracle(cases.riemann(8))


# --------------------------------------------------------------------------
# A2 equilibria: the discrete moments must equal the Maxwellian moments the
# model is built to reproduce (reference inlineFunctions.cpp:10-61,
# cpuLBM.cpp:785-844).
# --------------------------------------------------------------------------
def test_equilibrium_moment_identities(si_oracle):
    o = si_oracle
    k = o.consts()
    R, Tc, K, c = o.case.params["R"], k["Tc"], k["K"], k["c"]
    rng = np.random.default_rng(0)
    for _ in range(300):
        rho = rng.uniform(0.05, 4.0)
        U, V = rng.uniform(-0.7, 0.7, 2) * c
        T = rng.uniform(0.05, 0.8) * Tc
        f, h, failed = o.equilibrium(rho, U, V, T)
        assert failed == 0
        u, v, t = U / c, V / c, T / Tc
        s = rho * R * Tc
        pairs = [
            (f.sum(), rho),
            ((f * EX).sum(), rho * u), ((f * EY).sum(), rho * v),
            ((f * EX * EX).sum(), rho * (u * u + t)), ((f * EY * EY).sum(), rho * (v * v + t)),
            ((f * EX * EY).sum(), rho * u * v),
            ((f * EX ** 3).sum(), rho * (u ** 3 + 3 * u * t)),
            ((f * EX * EX * EY).sum(), rho * (u * u * v + v * t)),
            ((f * EY ** 3).sum(), rho * (v ** 3 + 3 * v * t)),
            (h.sum() / s, (u * u + v * v + (K + 2) * t) / 2 * rho / rho),
            ((h * EX).sum() / s, u * (u * u + v * v + (K + 4) * t) / 2),
            ((h * EY).sum() / s, v * (u * u + v * v + (K + 4) * t) / 2),
        ]
        for a, b in pairs:
            assert abs(a - b) <= 1e-13 * max(1.0, abs(b))


def test_equilibrium_symmetries(si_oracle):
    """x<->y exchange maps directions 1<->2, 3<->4, 6<->8, 9<->10, 11<->12;
    u -> -u maps 1<->3, 5<->6, 7<->8, 9<->11 (reference argument
    permutations, cpuLBM.cpp:795-823)."""
    o = si_oracle
    c, Tc = o.consts()["c"], o.consts()["Tc"]
    swap = [0, 2, 1, 4, 3, 5, 8, 7, 6, 10, 9, 12, 11]
    flipx = [0, 3, 2, 1, 4, 6, 5, 8, 7, 11, 10, 9, 12]
    rng = np.random.default_rng(1)
    for _ in range(50):
        rho, T = rng.uniform(0.2, 2), rng.uniform(0.1, 0.5) * Tc
        U, V = rng.uniform(-0.5, 0.5, 2) * c
        f, h, _ = o.equilibrium(rho, U, V, T)
        fs, hs, _ = o.equilibrium(rho, V, U, T)
        fx, hx, _ = o.equilibrium(rho, -U, V, T)
        np.testing.assert_allclose(fs, f[swap], rtol=1e-13, atol=1e-15)
        np.testing.assert_allclose(hs, h[swap], rtol=1e-13, atol=1e-9)
        np.testing.assert_allclose(fx, f[flipx], rtol=1e-13, atol=1e-15)
        np.testing.assert_allclose(hx, h[flipx], rtol=1e-13, atol=1e-9)


def test_equilibrium_mass_check_trips():
    """|sum feq - rho| > 1e-6 sets the failure flag (cpuLBM.cpp:825-843):
    at rho = 1e11 the rounding of the sum exceeds 1e-6; NaN never trips
    it (fabs(NaN) > x is false)."""
    o = Oracle(cases.couette(8))
    _, _, failed = o.equilibrium(1e11, 0.01, 0.0, 1.0)
    assert failed == 1
    _, _, failed = o.equilibrium(1.0, 0.01, 0.0, 1.0)
    assert failed == 0
    _, _, failed = o.equilibrium(float("nan"), 0.0, 0.0, 1.0)
    assert failed == 0


# --------------------------------------------------------------------------
# A3 moments: initialisation from the property table round-trips.
# --------------------------------------------------------------------------
def test_init_round_trip_macros():
    c = cases.riemann(16)
    o = Oracle(c)
    m = o.get_macro()
    R = c.params["R"]
    for pid in range(len(c.props["density"])):
        sel = c.bnd == pid
        np.testing.assert_allclose(m["rho"][sel], c.props["density"][pid], rtol=1e-14)
        np.testing.assert_allclose(m["ux"][sel], c.props["velocity_x"][pid], rtol=1e-12, atol=1e-10)
        np.testing.assert_allclose(m["uy"][sel], c.props["velocity_y"][pid], rtol=1e-12, atol=1e-10)
        np.testing.assert_allclose(m["T"][sel], c.props["temp"][pid], rtol=1e-12)
        np.testing.assert_allclose(m["p"][sel], m["rho"][sel] * R * m["T"][sel], rtol=1e-15)


# --------------------------------------------------------------------------
# A4 index / mirror maps: brute force against the definitions.
# --------------------------------------------------------------------------
@pytest.mark.parametrize("mode", ["periodic", "mirror", "clamp"])
def test_index_map_brute_force(mode):
    n = 5
    c = cases.riemann(n)
    for k in ("periodic_x", "periodic_y", "mirror_north", "mirror_south", "mirror_east", "mirror_west"):
        c.params[k] = 0
    if mode == "periodic":
        c.params["periodic_x"] = c.params["periodic_y"] = 1
    elif mode == "mirror":
        for k in ("mirror_north", "mirror_south", "mirror_east", "mirror_west"):
            c.params[k] = 1
    o = Oracle(c)
    for d in (0, 1):
        for i in range(n):
            for pm in range(-3, 4):
                j = i + pm
                r, mir = o.index(i, pm, d)
                if 0 <= j < n:
                    assert (r, mir) == (j, 0)
                elif mode == "periodic":
                    assert (r, mir) == (j % n, 0)
                elif mode == "mirror":
                    exp = -j - 1 if j < 0 else 2 * n - j - 1
                    assert (r, mir) == (exp, 1)
                else:
                    assert (r, mir) == (0 if j < 0 else n - 1, 0)


def test_mirror_direction_maps():
    """Reflection about x negates e_y; about y negates e_x."""
    for i in range(13):
        ix, iy = pyoracle.mirror_dirs(i, 1, 1)
        assert (EX[ix], EY[ix]) == (EX[i], -EY[i])
        assert (EX[iy], EY[iy]) == (-EX[i], EY[i])
        assert pyoracle.mirror_dirs(i, 0, 0) == (i, i)


# --------------------------------------------------------------------------
# A5 stencil: a solid row is replaced by linear extrapolation of the two
# nearest fluid points, and the point beyond it by the same value.
# --------------------------------------------------------------------------
def test_stencil_solid_extrapolation():
    c = cases.couette(10, nx=6)          # walls at rows 0 and 9, periodic x
    o = Oracle(c)
    rng = np.random.default_rng(3)
    f = rng.uniform(0.5, 1.5, (6, 10, 13))
    # cell (2, 1), direction 2 (e_y > 0): y-stencil rows -1(clamp->0), 0, 1, 2, 3
    Sx, Sy = o.stencil(f, 2, 1, 2, 5)
    ext = 2 * f[2, 1, 2] - f[2, 2, 2]
    np.testing.assert_allclose(Sy, [ext, ext, f[2, 1, 2], f[2, 2, 2], f[2, 3, 2]], rtol=0, atol=0)
    # same cell, direction 4 (e_y < 0): rows 3, 2, 1, 0, -1(clamp 0)
    Sx, Sy = o.stencil(f, 2, 1, 4, 5)
    ext = 2 * f[2, 1, 4] - f[2, 2, 4]
    np.testing.assert_allclose(Sy, [f[2, 3, 4], f[2, 2, 4], f[2, 1, 4], ext, ext], rtol=0, atol=0)
    # cell (2, 2), direction 2: rows 0(solid), 1, 2, 3, 4
    Sx, Sy = o.stencil(f, 2, 2, 2, 5)
    np.testing.assert_allclose(Sy, [2 * f[2, 1, 2] - f[2, 2, 2], f[2, 1, 2], f[2, 2, 2], f[2, 3, 2], f[2, 4, 2]])
    # x-stencil with periodic wrap, direction 1 at x = 0: columns 4, 5, 0, 1, 2
    Sx, Sy = o.stencil(f, 0, 4, 1, 5)
    np.testing.assert_allclose(Sx, f[[4, 5, 0, 1, 2], 4, 1])
    # NND (3 points) next to the wall
    Sx, Sy = o.stencil(f, 2, 1, 2, 3)
    np.testing.assert_allclose(Sy, [2 * f[2, 1, 2] - f[2, 2, 2], f[2, 1, 2], f[2, 2, 2]])


def test_stencil_mirror_remaps_direction():
    c = cases.riemann(6)
    c.params["mirror_west"] = 1
    o = Oracle(c)
    rng = np.random.default_rng(4)
    f = rng.uniform(0.5, 1.5, (6, 6, 13))
    # direction 1 at x = 0: columns -2 -> 1 (dir 3), -1 -> 0 (dir 3), 0, 1, 2 (dir 1)
    Sx, _ = o.stencil(f, 0, 3, 1, 5)
    np.testing.assert_allclose(Sx, [f[1, 3, 3], f[0, 3, 3], f[0, 3, 1], f[1, 3, 1], f[2, 3, 1]])


# --------------------------------------------------------------------------
# A6 WENO5 and A7 NND: known answers.
# --------------------------------------------------------------------------
def test_weno5_known_answers():
    w = pyoracle.weno5
    assert w(0.0, [1, 2, 3, 4, 5]) == 0.0
    for e in (0.7, -1.3, 466.0):
        v = abs(e)
        assert w(e, [2.5] * 5) == pytest.approx(v * 2.5, rel=1e-15)
        # linear data: all three candidate stencils agree -> exact face value
        S = 1.0 + 0.3 * np.arange(5)
        assert w(e, S) == pytest.approx(v * (S[2] + 0.5 * 0.3), rel=1e-14)
    # smooth data: weights -> linear weights (0.1, 0.6, 0.3): the
    # fifth-order upwind face value (2S0 - 13S1 + 47S2 + 27S3 - 3S4)/60
    x = 0.01 * (np.arange(5) - 2)
    S = 1.0 + np.sin(x) * 50
    lin = (2 * S[0] - 13 * S[1] + 47 * S[2] + 27 * S[3] - 3 * S[4]) / 60
    assert w(1.0, S) == pytest.approx(lin, rel=1e-9)
    # a jump between S2 and S3: the weight moves onto the left substencil
    S = np.array([1.0, 1.0, 1.0, 5.0, 5.0])
    assert w(1.0, S) == pytest.approx(1.0, rel=1e-8)
    S = np.array([1.0, 1.0, 5.0, 5.0, 5.0])
    assert w(1.0, S) == pytest.approx(5.0, rel=1e-8)


def test_minmod_and_nnd():
    mm = pyoracle.minmod
    assert mm(1.0, 2.0) == 1.0 and mm(2.0, 1.0) == 1.0
    assert mm(-1.0, -3.0) == -1.0 and mm(1.0, -1.0) == 0.0 and mm(0.0, 2.0) == 0.0
    assert mm(2.0, 2.0) == 2.0
    n = pyoracle.nnd
    assert n(0.0, [1, 2, 3]) == 0.0
    assert n(-2.0, [3, 3, 3]) == 6.0
    # monotone linear: exact face value v*(S1 + slope/2)
    assert n(1.5, [1.0, 2.0, 3.0]) == pytest.approx(1.5 * 2.5)
    # local extremum: first order (minmod = 0)
    assert n(1.0, [1.0, 3.0, 1.0]) == 3.0


# --------------------------------------------------------------------------
# A8/A10 fluxes: telescoping on a periodic grid and the uniform fixed point.
# --------------------------------------------------------------------------
def test_net_flux_telescopes_on_periodic_grid():
    c = cases.shear_wave(12)
    o = Oracle(c)
    rng = np.random.default_rng(5)
    f = 1.0 + 0.2 * rng.standard_normal((12, 12, 13))
    h = 2.0 + 0.2 * rng.standard_normal((12, 12, 13))
    ff, hf = o.boundary_fluxes(f, h)
    for i in range(13):
        tot_f = sum(o.cflux(ff, x, y, i) for x in range(12) for y in range(12))
        tot_h = sum(o.cflux(hf, x, y, i) for x in range(12) for y in range(12))
        scale = np.abs(ff[..., i, :]).sum() + 1e-300
        assert abs(tot_f) <= 1e-13 * scale
        assert abs(tot_h) <= 1e-13 * (np.abs(hf[..., i, :]).sum() + 1e-300)


def test_uniform_equilibrium_is_fixed_point():
    c = cases.shear_wave(8, u0=0.0)
    o = Oracle(c)
    f0, h0 = o.get_dist()
    o.step(3)
    f1, h1 = o.get_dist()
    np.testing.assert_allclose(f1, f0, rtol=1e-14, atol=0)
    np.testing.assert_allclose(h1, h0, rtol=1e-14, atol=0)


# --------------------------------------------------------------------------
# A9 walls: mass balance and conservation.
# --------------------------------------------------------------------------
def test_wall_closure_balances_mass():
    c = cases.couette(12, nx=6, u_wall=0.2)
    o = Oracle(c)
    rng = np.random.default_rng(6)
    f, h = o.get_dist()
    f = f * (1 + 0.05 * rng.standard_normal(f.shape))
    ff, _ = o.boundary_fluxes(f, h)
    dx = o.consts()["dx"]
    # bottom wall row 0: in-flux from the fluid row above through directions
    # pointing down equals the alpha-scaled emission back up
    for x in range(6):
        inflow = sum(ff[x, 1, i, 1] for i in (4, 7, 8, 12)) / dx
        outflow = sum(ff[x, 0, i, 1] for i in (2, 6, 5, 10)) / dx
        assert outflow == pytest.approx(inflow, rel=1e-13)


def test_couette_conserves_mass():
    c = cases.couette(16, nx=8, u_wall=0.15)
    o = Oracle(c)
    fluid = c.bnd == 0
    m0 = o.get_macro()["rho"][fluid].sum()
    o.step(40)
    m1 = o.get_macro()["rho"][fluid].sum()
    assert abs(m1 - m0) <= 1e-13 * m0


def test_periodic_energy_conserved():
    c = cases.shear_wave(12, u0=0.1)
    o = Oracle(c)
    _, h0 = o.get_dist()
    f0, _ = o.get_dist()
    o.step(20)
    f1, h1 = o.get_dist()
    assert abs(h1.sum() - h0.sum()) <= 1e-13 * abs(h0.sum())
    assert abs(f1.sum() - f0.sum()) <= 1e-13 * f0.sum()
    for e in (EX, EY):
        assert abs((f1 * e).sum() - (f0 * e).sum()) <= 1e-13 * np.abs(f0 * e).sum()


# --------------------------------------------------------------------------
# A11 stepper: IMEX-SSP2(3,2,2) relaxation factor and its order.
# --------------------------------------------------------------------------
def _relaxation_factor(a: float) -> float:
    g1 = 1 / (1 + a)
    g2 = (1 + a * g1) / (1 + a)
    g3 = (1 - a * g2) / (1 + a)
    return 1 - a * (g2 + g3)


def test_pure_relaxation_closed_form():
    """Spatially uniform state with a non-equilibrium part that carries no
    mass or momentum, u = 0: fluxes vanish, the equilibrium is fixed and each
    step multiplies f - feq by R(dt/(2 tau))."""
    c = cases.shear_wave(6, u0=0.0, tau=0.05)
    c.params["cfl"] = 0.3
    o = Oracle(c)
    f, h = o.get_dist()
    k = o.consts()
    # perturbation orthogonal to 1, e_x, e_y and with zero energy change
    p = np.zeros(13)
    p[9], p[11], p[10], p[12] = 1.0, 1.0, -1.0, -1.0  # +x2 pair up, +y2 pair down
    feq = f.copy()
    f2 = f + 1e-3 * p
    o.set_dist(f2, h)
    o.step(1)
    f3, _ = o.get_dist()
    tau = c.meta["tau"]
    a = k["dt"] / (2 * tau)
    ratio = (f3 - feq)[..., 9] / (f2 - feq)[..., 9]
    np.testing.assert_allclose(ratio, _relaxation_factor(a), rtol=1e-9)


@pytest.mark.parametrize("law,param,mu_expect", [
    # VHS (PY/codeCL.py:987): mu_ref (T/T_ref)^(0.5 + upsilon), T/T_ref = 1.25
    ("vhs", 0.25, lambda mu_ref: mu_ref * 1.25 ** 0.75),
    ("vhs", -0.5, lambda mu_ref: mu_ref),
    # Sutherland (PY/codeCL.py:989): mu_ref (T/T_ref)^1.5 (T_ref + S)/(T + S)
    ("sutherland", 0.4, lambda mu_ref: mu_ref * 1.25 ** 1.5 * (0.8 + 0.4) / (1.0 + 0.4)),
])
def test_viscosity_law_sets_relaxation(law, param, mu_expect):
    """The pure-relaxation state (T stays T0 = 1) relaxes with tau =
    mu(T0)/p for the temperature-dependent laws of the OpenCL front end:
    the decay ratio is R(dt/(2 tau)) with mu(T0) from the published law at
    T0/T_ref = 1.25 (a wrong exponent, an inverted ratio or the deck mu in
    its place would each miss it)."""
    c = cases.shear_wave(6, u0=0.0, tau=0.05)
    mu_ref = 0.03
    c.meta["viscosity"] = {"model": law, "mu_ref": mu_ref, "T_ref": 0.8, "param": param}
    o = Oracle(c)
    f, h = o.get_dist()
    k = o.consts()
    p = np.zeros(13)
    p[9], p[11], p[10], p[12] = 1.0, 1.0, -1.0, -1.0
    f2 = f + 1e-3 * p
    o.set_dist(f2, h)
    o.step(1)
    f3, _ = o.get_dist()
    tau = mu_expect(mu_ref) / 1.0      # p = rho R T = 1
    ratio = (f3 - f)[..., 9] / (f2 - f)[..., 9]
    np.testing.assert_allclose(ratio, _relaxation_factor(k["dt"] / (2 * tau)), rtol=1e-9)


def test_relaxation_factor_second_order():
    """R(a) must match exp(-2a) (exact decay over dt = 2a*tau) to O(a^3)."""
    errs = [abs(_relaxation_factor(a) - math.exp(-2 * a)) for a in (1e-2, 5e-3, 2.5e-3)]
    rates = [math.log(errs[i] / errs[i + 1], 2) for i in range(2)]
    assert all(2.7 < r < 3.3 for r in rates), rates


# --------------------------------------------------------------------------
# Whole trajectories: symmetry and the analytic Couette profile.
# --------------------------------------------------------------------------
def test_riemann_diagonal_symmetry():
    """The quadrant states are symmetric about the diagonal, so rho(x,y) =
    rho(y,x) and ux(x,y) = uy(y,x). Transposition permutes summation orders,
    and this shock problem amplifies rounding by ~1e3 per step, so the bar
    is 1e-9 relative: a sign or index slip gives O(1e-3)."""
    c = cases.riemann(24, steps=6)
    o = Oracle(c)
    o.step(6)
    m = o.get_macro()
    assert np.max(np.abs(m["rho"] - m["rho"].T)) <= 1e-9 * np.max(m["rho"])
    assert np.max(np.abs(m["ux"] - m["uy"].T)) <= 1e-9 * np.max(np.abs(m["ux"]))


def test_couette_steady_profile_matches_slip_solution():
    """Couette flow (30 fluid rows, Kn = 0.063) run to steady state: the
    velocity profile is the linear slip profile of couetteConstT.m.

    The steady solution of this scheme carries a period-4 standing wiggle
    (about 2% of u_wall in u and 5e-3 in T at this resolution, the same for
    any nx); the viscous-heating temperature rise (3e-4) is below it, so
    the temperature part of the closed form is not asserted."""
    c = cases.couette(32, nx=4, u_wall=0.1, tau=0.05, cfl=0.5)
    o = Oracle(c)
    k = o.consts()
    o.step(int(60.0 / k["dt"]))
    m = o.get_macro()
    ys = np.arange(1, c.ny - 1)
    u_sim = m["ux"][0, 1:c.ny - 1]
    u_ex, _, prm = cases.couette_analytic(c, ys)
    assert np.max(np.abs(u_sim - u_ex)) <= 0.03 * c.meta["u_wall"]
    slope = np.polyfit((ys - 0.5) / (c.ny - 2), u_sim, 1)[0]
    assert slope == pytest.approx(prm["A"], rel=0.05)
    # the slip length is visible: the profile does not reach the wall speeds
    assert u_sim[0] > 0.2 * prm["A"] / (c.ny - 2)


# --------------------------------------------------------------------------
# Non-uniform grid of the OpenCL front end (PY/codeCL.py:611-618, 676-679):
# face fluxes weighted by face length, net divided by cell area. The
# finite-volume form telescopes, so on a doubly periodic grid the sums
# weighted by the cell areas dx[x] dy[y] are conserved to round-off; a wrong
# index (the upwind cell's spacing, a swapped dx/dy) breaks that.
# --------------------------------------------------------------------------
def _weighted_totals(o, dx, dy):
    f, h = o.get_dist()
    vol = np.outer(dx, dy)
    return np.array([(vol * f.sum(axis=2)).sum(), (vol * (f * EX).sum(axis=2)).sum(),
                     (vol * (f * EY).sum(axis=2)).sum(), (vol * h.sum(axis=2)).sum()])


def test_spacing_uniform_arrays_are_the_scalar_path():
    c = cases.shear_wave(24, steps=10)
    a = Oracle(c)
    a.step(10)
    fa, ha = a.get_dist()
    a.close()
    b = Oracle(cases.with_spacing(c, np.full(c.nx, c.params["dx"]), np.full(c.ny, c.params["dy"])))
    b.step(10)
    fb, hb = b.get_dist()
    b.close()
    assert np.array_equal(fa, fb) and np.array_equal(ha, hb)


def _modulate(o, nx, ny):
    """Make the state vary along both axes (the shear wave varies along y
    only, and an x-invariant field has zero x-flux differences)."""
    f, h = o.get_dist()
    m = 1 + 0.02 * np.outer(np.sin(2 * np.pi * np.arange(nx) / nx), np.cos(2 * np.pi * np.arange(ny) / ny))
    o.set_dist(f * m[:, :, None], h * m[:, :, None])


@pytest.mark.parametrize("rk,fm", [(1, 1), (0, 1), (2, 1), (3, 1), (1, 0)],
                         ids=["rk3", "rk1", "euler", "euler_eq", "rk3_nnd"])
def test_spacing_conserves_area_weighted_totals(rk, fm):
    c = cases.stretched(cases.shear_wave(32, u0=0.15, steps=40, rk_method=rk, flux_method=fm), amp=0.3)
    dx = np.array(c.meta["spacing"]["dx"])
    dy = np.array(c.meta["spacing"]["dy"])
    o = Oracle(c)
    _modulate(o, c.nx, c.ny)
    t0 = _weighted_totals(o, dx, dy)
    o.step(40)
    t1 = _weighted_totals(o, dx, dy)
    f_spaced, _ = o.get_dist()
    o.close()
    scale = np.array([t0[0], t0[0], t0[0], t0[3]])
    drift = np.abs(t1 - t0) / scale
    assert np.all(drift <= 1e-13), drift
    # and the spacing changes the flow (the unweighted mass is not conserved)
    u = Oracle(cases.shear_wave(32, u0=0.15, steps=40, rk_method=rk, flux_method=fm))
    _modulate(u, c.nx, c.ny)
    u.step(40)
    f_uniform, _ = u.get_dist()
    u.close()
    assert np.abs(f_spaced - f_uniform).max() > 1e-6
    start = Oracle(c)
    _modulate(start, c.nx, c.ny)
    w0 = start.get_dist()[0].sum()
    start.close()
    w = f_spaced.sum()
    assert abs(w - w0) / w > 1e-12


def _planted_oracle(tmp_path, plant, macro="ORC_PLANT_SPACING"):
    """The oracle built with a planted mistake (`macro` = `plant`), loaded
    through a private copy of the pyoracle module."""
    import importlib.util
    import subprocess
    here = os.path.dirname(pyoracle.__file__)
    so = tmp_path / f"liblbm_oracle_{macro}{plant}.so"
    subprocess.run(["gcc", "-O2", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-std=c11",
                    "-D_POSIX_C_SOURCE=199309L", f"-D{macro}={plant}", "-shared", "-o", str(so),
                    os.path.join(here, "lbm_oracle.c"), "-lm"], check=True)
    spec = importlib.util.spec_from_file_location(f"pyoracle_{macro}{plant}", pyoracle.__file__)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    mod.LIB_PATH = str(so)
    return mod


@pytest.mark.parametrize("plant", [1, 2], ids=["upwind_dx", "dx_dy_swapped"])
def test_spacing_pin_catches_planted_mistakes(tmp_path, plant):
    """The area-weighted conservation pin fails on an oracle built with a
    wrong spacing index in the net flux."""
    mod = _planted_oracle(tmp_path, plant)
    c = cases.stretched(cases.shear_wave(32, u0=0.15, steps=40), amp=0.3)
    dx = np.array(c.meta["spacing"]["dx"])
    dy = np.array(c.meta["spacing"]["dy"])
    o = mod.Oracle(c)
    _modulate(o, c.nx, c.ny)
    t0 = _weighted_totals(o, dx, dy)
    o.step(40)
    t1 = _weighted_totals(o, dx, dy)
    o.close()
    drift = np.abs(t1 - t0) / np.array([t0[0], t0[0], t0[0], t0[3]])
    assert drift.max() > 1e-9, drift


def test_spacing_time_step_and_refusals():
    c = cases.shear_wave(16, steps=5)
    dx = np.full(c.nx, c.params["dx"])
    dy = np.full(c.ny, c.params["dy"])
    dy[3] = 0.25 * c.params["dy"]
    o = Oracle(c)
    dt0 = o.consts()["dt"]
    o.set_spacing(dx, dy)
    assert o.consts()["dt"] == pytest.approx(0.25 * dt0, rel=1e-15)   # dt from the smallest spacing
    with pytest.raises(ValueError):
        o.set_spacing(dx, -dy)
    o.close()


# --------------------------------------------------------------------------
# Spacing with walls: the C++ wall closure (cpuLBM.cpp:1502-1590) with the
# OpenCL front end's face-length weights (PY/codeCL.py:676-679, 722-737),
# i.e. the wall cell's own dx, dy in its mass balance. Pinned by what the
# balance is for: with walls (and no permanent cells) the fluid's
# area-weighted mass sum rho dx dy changes only by the wall faces' net
# flux, which the balance makes zero, so it is conserved to rounding on a
# stretched grid; and with uniform arrays the scalar path is reproduced
# bit for bit.
# --------------------------------------------------------------------------
def _walls_weighted_mass(o, dx, dy, fluid):
    f, _ = o.get_dist()
    return float((np.outer(dx, dy) * f.sum(axis=2))[fluid].sum())


def _stretched_obstacle(rk=1, fm=1):
    """Flow past a 6 x 6 block on a stretched doubly periodic grid: the
    block's corner cells have fluid neighbours on both axes, where the
    wall balance's x and y weights do not cancel."""
    base = cases.obstacle(32, block=(12, 12, 6, 6), u0=0.15, steps=60, rk_method=rk, flux_method=fm)
    c = cases.stretched(base, amp=0.3)
    fluid = np.asarray(c.bnd) == 0
    return c, np.array(c.meta["spacing"]["dx"]), np.array(c.meta["spacing"]["dy"]), fluid


@pytest.mark.parametrize("rk,fm", [(1, 1), (0, 1), (1, 0)], ids=["rk3", "rk1", "rk3_nnd"])
def test_spacing_with_walls_conserves_area_weighted_mass(rk, fm):
    c, dx, dy, fluid = _stretched_obstacle(rk, fm)
    o = Oracle(c)
    _modulate(o, c.nx, c.ny)
    m0 = _walls_weighted_mass(o, dx, dy, fluid)
    assert o.step(60) == 0
    m1 = _walls_weighted_mass(o, dx, dy, fluid)
    f_spaced, _ = o.get_dist()
    o.close()
    print(f"stretched obstacle rk{rk} fm{fm}: weighted mass drift {abs(m1 - m0) / m0:.1e}")
    assert abs(m1 - m0) <= 1e-13 * m0
    # the spacing matters: the uniform deck ends elsewhere
    u = Oracle(cases.obstacle(32, block=(12, 12, 6, 6), u0=0.15, steps=60, rk_method=rk, flux_method=fm))
    _modulate(u, c.nx, c.ny)
    u.step(60)
    f_uniform, _ = u.get_dist()
    u.close()
    assert np.abs(f_spaced - f_uniform).max() > 1e-6


def test_spacing_with_walls_pin_catches_uniform_wall_balance(tmp_path):
    """An oracle whose wall balance keeps the deck's uniform dx, dy on the
    stretched grid leaks mass through the walls."""
    mod = _planted_oracle(tmp_path, 3)
    c, dx, dy, fluid = _stretched_obstacle()
    o = mod.Oracle(c)
    _modulate(o, c.nx, c.ny)
    m0 = _walls_weighted_mass(o, dx, dy, fluid)
    o.step(60)
    m1 = _walls_weighted_mass(o, dx, dy, fluid)
    o.close()
    assert abs(m1 - m0) > 1e-9 * m0, abs(m1 - m0) / m0


def test_spacing_with_walls_uniform_arrays_are_the_scalar_path():
    c = cases.obstacle(24, block=(8, 8, 5, 5), u0=0.15, steps=10)
    a = Oracle(c)
    a.step(10)
    fa, ha = a.get_dist()
    a.close()
    b = Oracle(cases.with_spacing(c, np.full(c.nx, c.params["dx"]), np.full(c.ny, c.params["dy"])))
    b.step(10)
    fb, hb = b.get_dist()
    b.close()
    assert np.array_equal(fa, fb) and np.array_equal(ha, hb)


# --------------------------------------------------------------------------
# A11 stepper, the flux half: with the relaxation switched off (mu -> inf,
# so dt/tau ~ 1e-14) IMEX-SSP2(3,2,2) reduces to its explicit part, Heun's
# method on the net flux: f* = f - F(f), f_new = f - (F(f) + F(f*))/2
# (reference RK3 stages, cpuLBM.cpp:1945, 1970, 2022, 2039). The Heun step is
# built here from the oracle's own face fluxes and cFLUX, which the tests
# above pin separately. And with the relaxation on, the coupled stepper is
# second order in time at fixed dx.
# --------------------------------------------------------------------------
def _net_flux(o, flux, nx, ny):
    out = np.zeros((nx, ny, 13))
    for x in range(nx):
        for y in range(ny):
            for i in range(13):
                out[x, y, i] = o.cflux(flux, x, y, i)
    return out


def _heun_gap(mod):
    c = cases.shear_wave(16, u0=0.15, tau=1e12, steps=1)
    o = mod.Oracle(c)
    f, h = o.get_dist()
    m = 1 + 0.05 * np.outer(np.sin(2 * np.pi * np.arange(16) / 16), np.cos(2 * np.pi * np.arange(16) / 16))
    f0, h0 = f * m[:, :, None], h * m[:, :, None]
    o.set_dist(f0, h0)
    ff, hf = o.boundary_fluxes(f0, h0)
    F1, H1 = _net_flux(o, ff, 16, 16), _net_flux(o, hf, 16, 16)
    ff2, hf2 = o.boundary_fluxes(f0 - F1, h0 - H1)
    F2, H2 = _net_flux(o, ff2, 16, 16), _net_flux(o, hf2, 16, 16)
    fh, hh = f0 - 0.5 * (F1 + F2), h0 - 0.5 * (H1 + H2)
    o.step(1)
    fo, ho = o.get_dist()
    o.close()
    return (float(np.linalg.norm(fo - fh) / np.linalg.norm(fh - f0)),
            float(np.linalg.norm(ho - hh) / np.linalg.norm(hh - h0)))


def test_stepper_without_relaxation_is_heun():
    gf, gh = _heun_gap(pyoracle)
    assert gf <= 1e-12 and gh <= 1e-12, (gf, gh)


@pytest.mark.parametrize("plant", [1, 2], ids=["final_weight", "stage3_drops_dF2"])
def test_heun_pin_catches_planted_flux_weights(tmp_path, plant):
    """Flux-weight mistakes that the conservation, fixed-point and symmetry
    pins let through: the final combination taking -dF2 instead of
    -(dF2 + dF3)/2, and stage 3 without dF2."""
    gf, gh = _heun_gap(_planted_oracle(tmp_path, plant, "ORC_PLANT_RK"))
    assert max(gf, gh) > 1e-3, (gf, gh)


def _time_order_ratios(mod):
    """Error ratios between successive halvings of dt (CFL 0.4 -> 0.05 at
    fixed dx) at t = 0.1, against a run at CFL 0.0125, on a doubly periodic
    wave modulated along both axes with tau = 0.05 (dt/tau up to 0.125)."""
    out = {}
    for cfl in (0.4, 0.2, 0.1, 0.05, 0.0125):
        c = cases.shear_wave(32, u0=0.15, tau=0.05, steps=1)
        c.params["cfl"] = cfl
        o = mod.Oracle(c)
        f, h = o.get_dist()
        m = 1 + 0.05 * np.outer(np.sin(2 * np.pi * np.arange(32) / 32), np.cos(2 * np.pi * np.arange(32) / 32))
        o.set_dist(f * m[:, :, None], h * m[:, :, None])
        o.step(int(round(0.1 / o.consts()["dt"])))
        out[cfl] = np.concatenate([a.ravel() for a in o.get_dist()])
        o.close()
    errs = [np.linalg.norm(out[k] - out[0.0125]) for k in (0.4, 0.2, 0.1, 0.05)]
    return [float(errs[i] / errs[i + 1]) for i in range(3)]


def test_stepper_second_order_in_time():
    r = _time_order_ratios(pyoracle)
    assert all(3.5 < x < 4.6 for x in r), r


@pytest.mark.parametrize("plant", [1, 2], ids=["final_weight", "stage3_drops_dF2"])
def test_time_order_catches_planted_flux_weights(tmp_path, plant):
    r = _time_order_ratios(_planted_oracle(tmp_path, plant, "ORC_PLANT_RK"))
    assert max(r) < 3.0, r


# --------------------------------------------------------------------------
# Temperature half of the thermal Couette closed form
# (Items/11.04.18 - Couette flow check for macrossan/couetteConstT.m:1-30):
# at steady state T = T0 + C s^2 + E across the gap, C = -mu A^2/(2k). With
# the lattice temperature of the reference deck (Tc = 2.5 T) the profile is
# free of the period-4 standing wiggle the scheme shows at Tc = T.
# --------------------------------------------------------------------------
def _couette_temperature(mod, u_wall):
    c = cases.couette(32, nx=4, u_wall=u_wall, tau=0.05, cfl=0.5)
    c.params["Tc_Tref"] = 2.5
    o = mod.Oracle(c)
    st = o.step(int(60.0 / o.consts()["dt"]))
    T = o.get_macro()["T"][0, 1:c.ny - 1]
    o.close()
    ys = np.arange(1, c.ny - 1)
    _, _, prm = cases.couette_analytic(c, ys)
    s = (ys - 0.5) / (c.ny - 2) - 0.5
    p = np.polyfit(s, T, 2)
    rise = T.max() - T.min()
    return dict(status=st, curvature=float(p[0]), C=prm["C"],
                fit_residual=float(np.sqrt(np.mean((T - np.polyval(p, s)) ** 2)) / rise),
                asymmetry=float(np.max(np.abs(T - T[::-1])) / rise))


def _temperature_checks(mod):
    a, b = _couette_temperature(mod, 0.1), _couette_temperature(mod, 0.2)
    return a, b, [a["status"] == 0 and b["status"] == 0,
                  abs(a["curvature"] / a["C"] - 1) <= 0.05 and abs(b["curvature"] / b["C"] - 1) <= 0.05,
                  a["fit_residual"] <= 0.02 and b["fit_residual"] <= 0.02,
                  a["asymmetry"] <= 0.05 and b["asymmetry"] <= 0.05,
                  3.8 <= b["curvature"] / a["curvature"] <= 4.2]


def test_couette_temperature_profile_is_quadratic():
    """Steady viscous heating (u_wall 0.1 and 0.2, Kn = 0.063): quadratic in
    s within 2 % of the rise, symmetric about the mid-plane, curvature within
    5 % of the closed form's C (measured 2 %), and scaling with u_wall^2."""
    a, b, ok = _temperature_checks(pyoracle)
    assert all(ok), (a, b, ok)


@pytest.mark.parametrize("plant", [3, 1], ids=["coupling_sign", "final_weight"])
def test_temperature_pin_catches_planted_mistakes(tmp_path, plant):
    a, b, ok = _temperature_checks(_planted_oracle(tmp_path, plant, "ORC_PLANT_RK"))
    assert not all(ok), (a, b)


# --------------------------------------------------------------------------
# A12 residual (runOne, cpuLBM.cpp:2322-2331): RMS over all nx*ny cells of
# the change of T over the step.
# --------------------------------------------------------------------------
def _residual_gap(mod):
    c = cases.couette(12, nx=10, u_wall=0.15, steps=5)
    o = mod.Oracle(c)
    o.step(3)
    T0 = o.get_macro()["T"]
    st, r = o.step(1, want_res=True)
    T1 = o.get_macro()["T"]
    o.close()
    expect = float(np.sqrt(np.mean((T1 - T0) ** 2)))
    return st, r, expect


def test_residual_is_rms_temperature_change():
    st, r, expect = _residual_gap(pyoracle)
    assert st == 0 and r > 0
    assert r == pytest.approx(expect, rel=1e-12)


def test_residual_pin_catches_planted_mean(tmp_path):
    st, r, expect = _residual_gap(_planted_oracle(tmp_path, 1, "ORC_PLANT_RES"))
    assert r != pytest.approx(expect, rel=1e-6)
