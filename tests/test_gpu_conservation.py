"""Conserved totals on the GPU (lbm_get_totals: FP64 wavefront-shuffle
reductions over the fluid cells): they equal the sums of the distributions
read back, and mass, momentum and energy are conserved on a doubly periodic
shear wave — to round-off (1e-12 relative) in FP64, and to 1e-6 in mixed
precision, whose FP64 moment planes change only by the telescoping net
fluxes (the FP32 stored deviations never enter the totals). Mass is also
checked with the diffuse walls of the Couette deck."""
import numpy as np
import pytest

from uq_amd import cases

pytestmark = pytest.mark.gpu

EX = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1, 2, 0, -2, 0], dtype=np.float64)
EY = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1, 0, 2, 0, -2], dtype=np.float64)


def test_totals_equal_sums_of_readback(gpu_lib):
    c = cases.couette(37, nx=45, u_wall=0.15)
    s = gpu_lib.Solver(c, "fp64")
    try:
        s.step(7)
        t = s.totals()
        f, h = s.get_dist()
        cc = s.info()["c"]
    finally:
        s.close()
    fluid = c.bnd == 0   # property 0 is the fluid row of the Couette deck
    ff, hh = f[fluid], h[fluid]
    ref = np.array([ff.sum(), (ff * EX).sum() * cc, (ff * EY).sum() * cc, hh.sum()])
    scale = np.array([ff.sum(), np.abs(ff * EX).sum() * cc, np.abs(ff * EY).sum() * cc, hh.sum()])
    print(f"[totals] gpu {t} numpy {ref}", flush=True)
    assert np.all(np.abs(t - ref) <= 1e-13 * scale)


@pytest.mark.parametrize("precision,bar", [("fp64", [1e-12] * 4), ("mixed", [1e-6] * 4)])
def test_periodic_conservation(gpu_lib, precision, bar):
    c = cases.shear_wave(64, u0=0.15, steps=300)
    s = gpu_lib.Solver(c, precision)
    try:
        t0 = s.totals()
        s.step(300)
        t1 = s.totals()
        cc = s.info()["c"]
    finally:
        s.close()
    drift = np.abs(t1 - t0) / np.array([t0[0], t0[0] * cc, t0[0] * cc, t0[3]])
    print(f"[totals] shear64 {precision}: relative drift mass {drift[0]:.1e} px {drift[1]:.1e} "
          f"py {drift[2]:.1e} energy {drift[3]:.1e}", flush=True)
    assert np.all(drift <= np.array(bar))


def test_couette_mass_conserved_with_walls(gpu_lib):
    c = cases.couette(64, u_wall=0.1, steps=400)
    s = gpu_lib.Solver(c, "fp64")
    try:
        m0 = s.totals()[0]
        s.step(400)
        m1 = s.totals()[0]
    finally:
        s.close()
    print(f"[totals] couette64 mass drift {abs(m1 - m0) / m0:.1e}", flush=True)
    assert abs(m1 - m0) <= 1e-12 * m0
