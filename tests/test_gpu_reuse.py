"""Device buffers of destroyed solvers are kept and handed to later solvers
of the same sizes (lbm_release_cached returns them): a solver created on
recycled buffers, which hold what the previous solver left in them, steps
bit for bit as one created on fresh buffers, alone and in a batch (the
served-stream tests cover requests admitted on the buffers of retired
ones)."""
import numpy as np
import pytest

from uq_amd import cases

pytestmark = pytest.mark.gpu


def _run(lbm, case, steps, prec="fp64"):
    s = lbm.Solver(case, prec)
    try:
        assert s.step(steps) == 0
        return s.get_dist()
    finally:
        s.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_recycled_buffers_step_as_fresh_ones(gpu_lib, prec):
    a = cases.couette(64, steps=20)
    b = cases.shear_wave(64, u0=0.15, steps=20)  # same sizes: its buffers go to the next a
    gpu_lib.release_cached()
    fa, ha = _run(gpu_lib, a, 20, prec)
    _run(gpu_lib, b, 20, prec)
    fa2, ha2 = _run(gpu_lib, a, 20, prec)
    assert np.array_equal(fa, fa2) and np.array_equal(ha, ha2)


def test_recycled_buffers_in_a_batch(gpu_lib):
    decks = [cases.couette(64, steps=10), cases.shear_wave(96, steps=10), cases.couette(64, u_wall=0.05, steps=10)]
    gpu_lib.release_cached()
    alone = [_run(gpu_lib, c, 10) for c in decks]
    for _ in range(2):  # the second pass runs entirely on recycled buffers
        solvers = [gpu_lib.Solver(c, "fp64") for c in decks]
        b = gpu_lib.Batch(solvers)
        try:
            assert b.step(10) == 0
            for s, (f, h) in zip(solvers, alone):
                f2, h2 = s.get_dist()
                assert np.array_equal(f, f2) and np.array_equal(h, h2)
        finally:
            b.close()
            for s in solvers:
                s.close()
    gpu_lib.release_cached()
