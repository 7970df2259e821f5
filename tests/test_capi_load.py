"""The C-ABI library loads, exports every entry point include/uqlbm.h
declares, agrees with the ctypes struct mirrors, and validates decks on the
host (CPU only: no compute call needs a GPU here)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from uq_amd import capi, cases, lbm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "uqlbm.h")


def _declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lbm_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = lbm.lib()
    names = _declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_struct_layouts_match_header(tmp_path):
    src = tmp_path / "sizes.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "uqlbm.h"\n'
        "int main(void){printf(\"%zu %zu %zu %zu %zu %zu %zu\\n\", sizeof(lbm_params), sizeof(lbm_props),"
        " sizeof(lbm_info), sizeof(lbm_timing), sizeof(lbm_request_result), offsetof(lbm_params, Tc_Tref),"
        " offsetof(lbm_info, device_bytes)); return 0;}\n")
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    got = [int(v) for v in out]
    exp = [C.sizeof(capi.LbmParams), C.sizeof(capi.LbmProps), C.sizeof(capi.LbmInfo),
           C.sizeof(capi.LbmTiming), C.sizeof(capi.LbmRequestResult),
           capi.LbmParams.Tc_Tref.offset, capi.LbmInfo.device_bytes.offset]
    assert got == exp


def _create_status(case, precision=0):
    d = capi.PackedDeck(case.params, case.bnd, case.props)
    h = C.c_void_p()
    st = lbm.lib().lbm_create(C.byref(d.params), case.nx, case.ny, d.bnd_ptr(), C.byref(d.props),
                              precision, C.byref(h))
    if st == 0:
        lbm.lib().lbm_destroy(h)
    return st, lbm.lib().lbm_last_error().decode()


@pytest.mark.parametrize("case", [
    cases.couette(64), cases.couette(37, nx=45), cases.shear_wave(40), cases.riemann(64),
    cases.flat_plate(80, 40, plate_start=20), cases.flat_plate(512, 256),
], ids=lambda c: c.name)
def test_create_accepts_workload_geometries(case):
    st, msg = _create_status(case)
    assert st in (capi.LBM_OK, capi.LBM_ERR_NOGPU), msg


@pytest.mark.parametrize("case", [
    cases.obstacle(48, block=(18, 20, 8, 1)), cases.obstacle(48, block=(18, 20, 1, 8)),
    cases.obstacle(48, block=(18, 1, 8, 3), mirror_south=True),
    cases.obstacle(48, block=(18, 2, 8, 1), mirror_south=True),
], ids=lambda c: c.name + ("_mirror" if c.params["mirror_south"] else ""))
def test_create_accepts_thin_and_mirror_edge_solids(case):
    """One-cell-thick plates and solids near a mirror edge, where the tile
    kernels' logical stencil addressing differs from the reference's
    physical addressing: accepted, their tiles flagged for the physical
    path (the GPU parity tests compare them with the oracle)."""
    st, msg = _create_status(case)
    assert st in (capi.LBM_OK, capi.LBM_ERR_NOGPU), msg


@pytest.mark.parametrize("mutate,expected", [
    (lambda c: c.params.__setitem__("rk_method", 7), capi.LBM_ERR_ARG),
    (lambda c: c.params.__setitem__("flux_method", 3), capi.LBM_ERR_ARG),
    (lambda c: c.bnd.__setitem__((0, 0), 99), capi.LBM_ERR_ARG),
    (lambda c: c.params.__setitem__("mu", 0.0), capi.LBM_ERR_ARG),
    (lambda c: c.params.__setitem__("Pr", -0.7), capi.LBM_ERR_ARG),
])
def test_create_rejects_bad_decks(mutate, expected):
    c = cases.couette(16)
    c.bnd = c.bnd.copy()
    mutate(c)
    st, msg = _create_status(c)
    assert st == expected, msg


def test_create_rejects_tiny_grid_and_bad_precision():
    st, _ = _create_status(cases.couette(8, nx=3))
    assert st == capi.LBM_ERR_ARG
    st, _ = _create_status(cases.couette(16), precision=5)
    assert st == capi.LBM_ERR_ARG


def test_null_handles_are_rejected():
    L = lbm.lib()
    assert L.lbm_init(None) == capi.LBM_ERR_ARG
    assert L.lbm_step(None, 1, 0, None) == capi.LBM_ERR_ARG
    assert L.lbm_batch_step(None, 1) == capi.LBM_ERR_ARG


def test_case_builders_are_consistent():
    ens = cases.uq_ensemble(32, seed=5, steps=100)
    assert len(ens) == 32
    assert {c.nx for c in ens} <= {64, 128, 256, 512}
    again = cases.uq_ensemble(32, seed=5, steps=100)
    assert [c.meta["tau"] for c in ens] == [c.meta["tau"] for c in again]
    for c in ens:
        assert 0.05 <= c.meta["u_wall"] <= 0.2 and 0.1 <= c.meta["mach"] <= 0.5
        T = c.meta["T"]
        assert T == pytest.approx(c.meta["u_wall"] ** 2 / (1.4 * c.meta["mach"] ** 2))
    assert all(c.params["Tc_Tref"] == 2.5 for c in ens)
    reqs = cases.request_stream(rate=50, duration=4.0, seed=1)
    assert 100 < len(reqs) < 320
    assert all(64 <= r.case.nx <= 1024 and 500 <= r.case.nsteps <= 5000 for r in reqs)
    assert all(reqs[k].arrival_s < reqs[k + 1].arrival_s for k in range(len(reqs) - 1))
    r = cases.riemann(16)
    assert np.array_equal(r.bnd, np.rot90(np.rot90(r.bnd)) * 0 + r.bnd)
    assert r.bnd[12, 12] == 0 and r.bnd[2, 12] == 1 and r.bnd[2, 2] == 2 and r.bnd[12, 2] == 3


def test_cpp_dropin_header_compiles(tmp_path):
    """include/gpuLBM.hpp compiles as plain C++ against the C ABI."""
    src = tmp_path / "use.cpp"
    src.write_text('#include "gpuLBM.hpp"\nint main(){ gpuLBM l; l.Nx = 4; (void)l; return 0; }\n')
    subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", os.path.join(ROOT, "include"), str(src)],
                   check=True)


def test_flux_shape_table():
    """The flux tile shapes the autotuner chooses from (host query, no GPU)."""
    dims = [lbm.flux_shape_dims(k) for k in range(lbm.FLUX_SHAPES)]
    assert dims == [(15, 16), (31, 15), (15, 32), (39, 12), (23, 20)]
    # every shape keeps one x face and one y face per thread of its block
    for (w, h), nt in zip(dims, (256, 512, 512, 512, 512)):
        assert (w + 1) * h <= nt and w * (h + 1) <= nt
    with pytest.raises(lbm.LbmError):
        lbm.flux_shape_dims(lbm.FLUX_SHAPES)


def test_strip_mode_switch():
    """lbm_set_strips takes modes 0 (tiles), 1 (size rule, the default) and
    2 (strips wherever possible) and refuses others (host setting, no GPU)."""
    try:
        for mode in (0, 2, 1):
            lbm.set_strips(mode)
        for bad in (-1, 3):
            with pytest.raises(lbm.LbmError):
                lbm.set_strips(bad)
    finally:
        lbm.set_strips(1)


def test_integration_binding_compiles(tmp_path):
    """The maintainer-side binding published with the repository (the
    `gpuStep.h` block of the integration notes) compiles against
    include/uqlbm.h, given a class with the public members
    of the reference's cpuLBM that it touches (declared here for the check
    only)."""
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    block = re.search(r"```cpp\n(// gpuStep\.h.*?)```", text, re.S).group(1)
    members = """
#include <cstdio>
struct double2 { double x, y; };
struct cpuLBM {
  int Nx, Ny; int* bndArray; int RKMETHOD, FMETHOD; double CFL, dtau, tt_tref, tt_time; int steps;
  double tol; int periodicX, periodicY, mirrorNorth, mirrorSouth, mirrorEast, mirrorWest;
  double gamma, R, mu, Pr, dx, dy, Lx, Ly, rho_ref, p_ref, T_ref, Tc_Tref;
  int numProps; double *density, *pressure, *temp, *velocityX, *velocityY; int* cellType;
  double *f_G, *h_G, *rho_G, *p_G, *T_G; double2* u_G; double residual; int step, nPrintOut, MAT;
};
"""
    src = tmp_path / "binding.cpp"
    src.write_text(members + block + "\nint main() { return 0; }\n")
    subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", os.path.join(ROOT, "include"), str(src)],
                   check=True)
