"""lbm_step_noalloc, the stepping entry point that allocates nothing (its
residual reduction uses a file-scope device array; kernels launched
directly): called through the C ABI on two solvers in a row, the second on
a smaller grid, it must give lbm_step's residual and state bit for bit; a
grid with more rows than the array serves is refused with LBM_ERR_ARG."""
import ctypes as C

import numpy as np
import pytest

from uq_amd import cases
from uq_amd.capi import LBM_ERR_ARG

pytestmark = pytest.mark.gpu


def _noalloc(lib):
    f = lib.lbm_step_noalloc
    f.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(C.c_double)]
    f.restype = C.c_int
    return f


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_step_equals_step(gpu_lib, precision):
    lib = gpu_lib.lib()
    step_noalloc = _noalloc(lib)
    for case, n in ((cases.couette(96, nx=80, u_wall=0.12, steps=30), 30),
                    (cases.riemann(40, steps=12), 12)):       # second call: a smaller grid
        a, b = gpu_lib.Solver(case, precision), gpu_lib.Solver(case, precision)
        try:
            st_a, r_a = a.step(n, want_residual=True)
            r_b = C.c_double(0.0)
            st_b = step_noalloc(b.handle, n, 1, C.byref(r_b))
            print(f"[noalloc] {case.name} {precision}: lbm_step residual {r_a:.17e}, "
                  f"lbm_step_noalloc {r_b.value:.17e}", flush=True)
            assert st_a == 0 and st_b == 0
            assert r_a == r_b.value
            fa, ha = a.get_dist()
            fb, hb = b.get_dist()
            assert np.array_equal(fa, fb) and np.array_equal(ha, hb)
            assert a.info()["step"] == b.info()["step"] == n
        finally:
            a.close()
            b.close()


def test_noalloc_refuses_too_many_rows(gpu_lib):
    lib = gpu_lib.lib()
    step_noalloc = _noalloc(lib)
    c = cases.shear_wave(4, steps=1)
    c = cases.Case("tall", dict(c.params), np.zeros((4, 32769), dtype=np.int32), c.props, 1, dict(c.meta))
    s = gpu_lib.Solver(c, "fp64")
    try:
        r = C.c_double(0.0)
        assert step_noalloc(s.handle, 1, 1, C.byref(r)) == LBM_ERR_ARG
        print(f"[noalloc] refused: {lib.lbm_last_error().decode()}", flush=True)
        assert step_noalloc(s.handle, 1, 0, None) == 0   # without the residual the array is not used
    finally:
        s.close()
