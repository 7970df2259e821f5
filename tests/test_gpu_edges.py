"""Edge sizes: the smallest grids the library accepts (4 cells a side, grids
narrower than one tile, walls two cells apart) against the oracle, and the
largest single instances this GPU holds checked through size-independent
properties (conservation, x invariance)."""
import numpy as np
import pytest

from uq_amd import cases
from uq_amd.lbm import rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", [
    cases.shear_wave(4, steps=10), cases.shear_wave(5, rk_method=0, steps=10), cases.couette(6, nx=4, steps=10),
    cases.couette(8, nx=5, steps=10), cases.riemann(8, steps=6), cases.flat_plate(16, 8, plate_start=4, steps=8),
], ids=lambda c: f"{c.name}_{c.nx}x{c.ny}_rk{c.params['rk_method']}")
def test_tiny_grids_match_oracle(gpu_lib, case):
    """Bar 1e-12, or 8x the oracle's own one-ulp sensitivity for the shock
    deck (the 8x8 Riemann problem: 2e-11 after 6 steps)."""
    from pyoracle import Oracle
    from test_gpu_parity import _oracle_floor
    bar = 1e-12
    if case.meta.get("kind") == "riemann" or case.name.startswith("riemann"):
        bar = max(bar, 8 * _oracle_floor(case, case.nsteps))
    print(f"[edges] {case.name} {case.nx}x{case.ny} bar {bar:.1e}", flush=True)
    o = Oracle(case)
    so = o.step(case.nsteps)
    fo, ho = o.get_dist()
    o.close()
    s = gpu_lib.Solver(case, "fp64")
    try:
        sg = s.step(case.nsteps)
        f, h = s.get_dist()
    finally:
        s.close()
    print(f"[edges] {case.name}: status {so}/{sg} f {rel_l2(f, fo):.1e} h {rel_l2(h, ho):.1e}", flush=True)
    assert so == sg == 0
    assert rel_l2(f, fo) <= bar and rel_l2(h, ho) <= bar


@pytest.mark.parametrize("n,precision", [(8192, "fp64"), (12288, "fp64")])
def test_large_single_instance(gpu_lib, n, precision):
    """A doubly periodic shear wave of n x n cells (12288^2 FP64: 151 M
    cells, 162 GB of device state): totals conserved to round-off over two
    steps, and at 8192^2 every column of u_x identical (x-invariance)."""
    c = cases.shear_wave(n, u0=0.1, steps=2)
    print(f"[edges] shear {n}^2 {precision}", flush=True)
    s = gpu_lib.Solver(c, precision)
    try:
        print(f"[edges] device bytes {s.info()['device_bytes'] / 1e9:.1f} GB", flush=True)
        t0 = s.totals()
        assert s.step(2) == 0
        t1 = s.totals()
        cc = s.info()["c"]
        drift = np.abs(t1 - t0) / np.array([t0[0], t0[0] * cc, t0[0] * cc, t0[3]])
        print(f"[edges] drift {drift}", flush=True)
        assert np.all(drift <= 1e-12)
        if n <= 8192:
            ux = s.get_macro()["ux"]
            assert np.all(ux == ux[:1, :])
    finally:
        s.close()
