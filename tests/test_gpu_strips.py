"""Strip flux kernels (one wavefront per 58-column strip of up to 32 rows,
marched row by row; lbm_set_strips) against the flux tile kernels they
replace (strips forced on: lbm_set_strips mode 2, since these decks are
smaller than the launch groups the strips take by default), on decks
covering every boundary treatment the strips take over:
doubly periodic (shear wave), periodic with solid walls (Couette, WENO5 and
NND), clamped edges (Riemann), a north/south mirror with a clamp opposite
and a solid block (obstacle), east/west mirrors, a stretched grid
(per-column dt/dx, per-row dt/dy), the one-stage steppers (RK1, Euler), grids
whose sides are no multiple of the strip width or height (taller than one
row block: the items of a strip column alternate between ascending and
descending row order, so both orders are compared), instances
narrower than one strip and shorter than one row block, and mixed
precision (FP32 storage: the stage-2 flux runs on strips, the
moment-summing stage-3 flux stays on tiles).

The two paths compute every face with the same arithmetic (WENO5, NND,
solid extrapolation, wall closure) and form the net fluxes and the stage-3
update with contraction off, so their results are bitwise equal: FP64, and
FP32 storage on decks near their reference equilibrium (both paths compute
their stencils in FP32 there). The oracle bar, max(1e-12, 8 x the oracle's
one-ulp sensitivity), is checked by the parity tests, which run on strips
by default."""
import numpy as np
import pytest

from uq_amd import cases
from uq_amd.cases import Case
from uq_amd.lbm import rel_l2

pytestmark = pytest.mark.gpu


def _east_west_mirrors(case: Case) -> Case:
    p = dict(case.params)
    p.update(periodic_x=0, mirror_west=1, mirror_east=1)
    return Case(case.name + "_ew_mirror", p, case.bnd, case.props, case.nsteps, dict(case.meta))


DECKS = {
    "shear_periodic": lambda: cases.shear_wave(96, u0=0.15, steps=30),
    "couette_walls": lambda: cases.couette(64, nx=100, steps=30),
    "couette_nnd": lambda: cases.couette(64, nx=70, flux_method=0, steps=30),
    "riemann_clamped": lambda: cases.riemann(128, steps=20, cfl=0.1),
    "obstacle_ns_mirror": lambda: cases.obstacle(96, block=(40, 40, 10, 8), u0=0.1, mirror_south=True, steps=30),
    "obstacle_ew_mirror": lambda: _east_west_mirrors(cases.obstacle(96, block=(40, 40, 10, 8), u0=0.1, steps=30)),
    "shear_stretched": lambda: cases.stretched(cases.shear_wave(80, u0=0.15, steps=30), amp=0.3),
    "shear_rk1": lambda: cases.shear_wave(64, u0=0.1, rk_method=0, steps=30),
    "shear_euler": lambda: cases.shear_wave(64, u0=0.1, rk_method=2, steps=30),
    # instances narrower than one strip and shorter than one row block: lanes
    # beyond the last column and rows beyond the last row map (and wrap) far
    # past the edges
    "tiny_couette": lambda: cases.couette(9, nx=13, steps=10),
    "tiny_shear": lambda: cases.shear_wave(12, u0=0.1, steps=10),
}


def _run(lbm, case, steps, prec, strips):
    lbm.set_strips(2 if strips else 0)
    try:
        s = lbm.Solver(case, prec)
        try:
            st = s.step(steps)
            f, h = s.get_dist()
            m = s.get_macro()
        finally:
            s.close()
    finally:
        lbm.set_strips(1)
    return st, f, h, m


@pytest.mark.parametrize("name", list(DECKS))
def test_strips_match_tiles_fp64(gpu_lib, name):
    case = DECKS[name]()
    steps = case.nsteps
    st_s, fs, hs, _ = _run(gpu_lib, case, steps, "fp64", True)
    st_t, ft, ht, _ = _run(gpu_lib, case, steps, "fp64", False)
    print(f"[strips] {name} {case.nx}x{case.ny} {steps} steps: f {rel_l2(fs, ft):.2e} h {rel_l2(hs, ht):.2e}",
          flush=True)
    assert st_s == 0 and st_t == 0
    assert np.array_equal(fs, ft) and np.array_equal(hs, ht)


@pytest.mark.parametrize("name", ["shear_periodic", "couette_walls", "obstacle_ns_mirror"])
def test_strips_match_tiles_mixed(gpu_lib, name):
    case = DECKS[name]()
    steps = case.nsteps
    st_s, fs, hs, _ = _run(gpu_lib, case, steps, "mixed", True)
    st_t, ft, ht, _ = _run(gpu_lib, case, steps, "mixed", False)
    print(f"[strips] {name} mixed {steps} steps: f {rel_l2(fs, ft):.2e} h {rel_l2(hs, ht):.2e}", flush=True)
    assert st_s == 0 and st_t == 0
    assert np.array_equal(fs, ft) and np.array_equal(hs, ht)


def test_strips_toggle_applies_to_later_steps(gpu_lib):
    """Turning strips on and off between step calls of one solver replays
    the right launch sequence (the captured graphs are keyed by it): the
    run equals one made with strips throughout."""
    case = DECKS["couette_walls"]()
    ref = _run(gpu_lib, case, 20, "fp64", True)
    s = gpu_lib.Solver(case, "fp64")
    try:
        for mode in (2, 0, 2, 0):
            gpu_lib.set_strips(mode)
            assert s.step(5) == 0
        f, h = s.get_dist()
    finally:
        gpu_lib.set_strips(1)
        s.close()
    assert np.array_equal(f, ref[1]) and np.array_equal(h, ref[2])
