"""ctypes mirrors of the POD structs declared in ``include/uqlbm.h``.

The field order and types here must match the header exactly; the
``tests/test_capi_load.py`` suite checks the sizes against the compiled
library's own ``sizeof`` reports.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

LBM_Q = 13

LBM_OK = 0
LBM_FAILED = 1
LBM_ERR_ARG = -1
LBM_ERR_HIP = -2
LBM_ERR_GEOMETRY = -3
LBM_ERR_NOMEM = -4
LBM_ERR_STATE = -5
LBM_ERR_NOGPU = -6

LBM_PREC_FP64 = 0
LBM_PREC_MIXED = 1

STATUS_NAMES = {
    0: "LBM_OK",
    1: "LBM_FAILED",
    -1: "LBM_ERR_ARG",
    -2: "LBM_ERR_HIP",
    -3: "LBM_ERR_GEOMETRY",
    -4: "LBM_ERR_NOMEM",
    -5: "LBM_ERR_STATE",
    -6: "LBM_ERR_NOGPU",
}


class LbmParams(C.Structure):
    """``lbm_params``: the reference deck scalars (cpuLBM.cpp:113-199)."""

    _fields_ = [
        ("rk_method", C.c_int32),
        ("flux_method", C.c_int32),
        ("cfl", C.c_double),
        ("dtau", C.c_double),
        ("tt_tref", C.c_double),
        ("tt_time", C.c_double),
        ("steps", C.c_int32),
        ("tol", C.c_double),
        ("periodic_x", C.c_int32),
        ("periodic_y", C.c_int32),
        ("mirror_north", C.c_int32),
        ("mirror_south", C.c_int32),
        ("mirror_east", C.c_int32),
        ("mirror_west", C.c_int32),
        ("gamma", C.c_double),
        ("R", C.c_double),
        ("mu", C.c_double),
        ("Pr", C.c_double),
        ("dx", C.c_double),
        ("dy", C.c_double),
        ("Lx", C.c_double),
        ("Ly", C.c_double),
        ("rho_ref", C.c_double),
        ("p_ref", C.c_double),
        ("T_ref", C.c_double),
        ("Tc_Tref", C.c_double),
    ]


class LbmProps(C.Structure):
    _fields_ = [
        ("num_props", C.c_int32),
        ("density", C.POINTER(C.c_double)),
        ("pressure", C.POINTER(C.c_double)),
        ("temp", C.POINTER(C.c_double)),
        ("velocity_x", C.POINTER(C.c_double)),
        ("velocity_y", C.POINTER(C.c_double)),
        ("cell_type", C.POINTER(C.c_int32)),
    ]


class LbmInfo(C.Structure):
    _fields_ = [
        ("nx", C.c_int32),
        ("ny", C.c_int32),
        ("nsteps", C.c_int32),
        ("step", C.c_int32),
        ("dt", C.c_double),
        ("c", C.c_double),
        ("Tc", C.c_double),
        ("b", C.c_double),
        ("K", C.c_double),
        ("dx", C.c_double),
        ("dy", C.c_double),
        ("tstep", C.c_double),
        ("precision", C.c_int32),
        ("halo", C.c_int32),
        ("device_bytes", C.c_size_t),
    ]


class LbmTiming(C.Structure):
    _fields_ = [
        ("total_ms", C.c_double),
        ("n_kernels", C.c_int32),
        ("names", (C.c_char * 32) * 8),
        ("avg_ms", C.c_double * 8),
        ("launches", C.c_int64 * 8),
    ]


class LbmRequestResult(C.Structure):
    _fields_ = [
        ("tag", C.c_int64),
        ("status", C.c_int32),
        ("steps_done", C.c_int32),
        ("submit_s", C.c_double),
        ("admit_s", C.c_double),
        ("finish_s", C.c_double),
        ("cells", C.c_int64),
        ("residual", C.c_double),
    ]


def dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def iptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


class PackedDeck:
    """Keeps the numpy buffers behind an ``lbm_params``/``lbm_props`` pair
    alive for as long as the ctypes structs are in use."""

    def __init__(self, params: dict, bnd: np.ndarray, props: dict):
        self.params = LbmParams()
        for name, _ in LbmParams._fields_:
            setattr(self.params, name, params.get(name, 0))
        self.bnd = np.ascontiguousarray(bnd, dtype=np.int32)
        self.nx, self.ny = self.bnd.shape
        self._arrays = {
            k: np.ascontiguousarray(props[k], dtype=np.float64)
            for k in ("density", "pressure", "temp", "velocity_x", "velocity_y")
        }
        self._arrays["cell_type"] = np.ascontiguousarray(props["cell_type"], dtype=np.int32)
        n = len(self._arrays["cell_type"])
        for k, v in self._arrays.items():
            if len(v) != n:
                raise ValueError(f"property table column {k} has {len(v)} rows, expected {n}")
        self.props = LbmProps(
            n,
            dptr(self._arrays["density"]),
            dptr(self._arrays["pressure"]),
            dptr(self._arrays["temp"]),
            dptr(self._arrays["velocity_x"]),
            dptr(self._arrays["velocity_y"]),
            iptr(self._arrays["cell_type"]),
        )

    def bnd_ptr(self):
        return iptr(self.bnd)
