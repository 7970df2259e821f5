"""Python binding of the C ABI in ``include/uqlbm.h`` (libuqlbm.so).

``Solver`` mirrors the reference's cpuLBM operations (create from a deck,
initialise, step, macroscopic and distribution output); ``Batch`` advances
many solvers in batched launches; ``Server`` is the continuous-batching
request loop. All compute runs in the gfx950 library; this module only
marshals arrays. Loading fails loudly when the library is missing.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .capi import (LBM_Q, LBM_FAILED, LBM_OK, LBM_PREC_FP64, LBM_PREC_MIXED, STATUS_NAMES,
                   LbmInfo, LbmRequestResult, LbmTiming, PackedDeck, dptr)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libuqlbm.so")

PRECISIONS = {"fp64": LBM_PREC_FP64, "mixed": LBM_PREC_MIXED}

_lib = None


class LbmError(RuntimeError):
    def __init__(self, status: int, where: str):
        self.status = status
        msg = lib().lbm_last_error().decode(errors="replace")
        super().__init__(f"{where}: {STATUS_NAMES.get(status, status)}: {msg}")


def lib():
    """Load libuqlbm.so (built by uq_amd.build); raise if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    vp, i32, d = C.c_void_p, C.c_int32, C.c_double
    L.lbm_last_error.restype = C.c_char_p
    L.lbm_device_count.restype = i32
    L.lbm_create.restype = C.c_int
    L.lbm_create.argtypes = [vp, i32, i32, P(i32), vp, i32, P(vp)]
    L.lbm_init.argtypes = [vp]
    L.lbm_step.argtypes = [vp, i32, i32, P(d)]
    L.lbm_get_macro.argtypes = [vp, P(d), P(d), P(d), P(d), P(d)]
    L.lbm_get_dist.argtypes = [vp, P(d), P(d)]
    L.lbm_set_dist.argtypes = [vp, P(d), P(d)]
    L.lbm_get_info.argtypes = [vp, P(LbmInfo)]
    L.lbm_destroy.argtypes = [vp]
    L.lbm_destroy.restype = None
    L.lbm_batch_create.argtypes = [P(vp), i32, P(vp)]
    L.lbm_batch_step.argtypes = [vp, i32]
    L.lbm_batch_destroy.argtypes = [vp]
    L.lbm_batch_destroy.restype = None
    L.lbm_set_kernel_timing.argtypes = [i32]
    L.lbm_set_graphs.argtypes = [i32]
    L.lbm_set_strips.argtypes = [i32]
    L.lbm_release_cached.argtypes = []
    L.lbm_release_cached.restype = None
    L.lbm_get_totals.argtypes = [vp, P(d)]
    L.lbm_set_flux_shape.argtypes = [i32, i32]
    L.lbm_set_viscosity.argtypes = [vp, i32, d, d, d]
    L.lbm_set_spacing.argtypes = [vp, P(d), P(d)]
    L.lbm_flux_shape_dims.argtypes = [i32, P(i32), P(i32)]
    L.lbm_batch_autotune.argtypes = [vp, i32, P(i32), P(d)]
    L.lbm_get_timing.argtypes = [P(LbmTiming)]
    L.lbm_clock_s.restype = d
    if hasattr(L, "lbm_server_create"):
        L.lbm_server_create.argtypes = [i32, i32, C.c_size_t, i32, P(vp)]
        L.lbm_server_submit.argtypes = [vp, vp, i32, i32, P(i32), vp, i32, C.c_int64]
        L.lbm_server_tick.argtypes = [vp, i32, P(i32), P(i32)]
        L.lbm_server_poll.argtypes = [vp, P(LbmRequestResult), i32]
        L.lbm_server_keep_results.argtypes = [vp, i32]
        L.lbm_server_result_macro.argtypes = [vp, C.c_int64, P(d), P(d), P(d), P(d), P(d)]
        L.lbm_server_stats.argtypes = [vp, P(C.c_int64), P(C.c_size_t)]
        L.lbm_server_destroy.argtypes = [vp]
        L.lbm_server_destroy.restype = None
    _lib = L
    return L


def device_count() -> int:
    return int(lib().lbm_device_count())


def _check(status: int, where: str, allow_failed: bool = False) -> int:
    if status == LBM_OK or (allow_failed and status == LBM_FAILED):
        return status
    raise LbmError(status, where)


def set_kernel_timing(enable: bool) -> None:
    lib().lbm_set_kernel_timing(1 if enable else 0)


FLUX_SHAPES = 5


def set_flux_shape(precision: str, shape: int) -> None:
    """Flux tile shape of solvers, batches and servers of `precision`
    created from now on (0: 15x16, 1: 31x15, 2: 15x32, 3: 39x12, 4: 23x20
    own cells)."""
    _check(lib().lbm_set_flux_shape(PRECISIONS[precision], int(shape)), "lbm_set_flux_shape")


def flux_shape_dims(shape: int) -> tuple[int, int]:
    w, h = C.c_int32(0), C.c_int32(0)
    _check(lib().lbm_flux_shape_dims(int(shape), C.byref(w), C.byref(h)), "lbm_flux_shape_dims")
    return w.value, h.value


def set_graphs(enable: bool) -> None:
    """Replay captured HIP graphs of the step sequence (default) or launch
    every kernel directly."""
    lib().lbm_set_graphs(1 if enable else 0)


def set_strips(mode) -> None:
    """Flux kernels of instances with logical addressing: 0 (or False) tiles
    only, 1 (or True; default) strips for launch groups large enough to fill
    the GPU, 2 strips wherever possible."""
    mode = int(mode)
    _check(lib().lbm_set_strips(mode), "lbm_set_strips")


def release_cached() -> None:
    """Return the device buffers kept from destroyed solvers to the device."""
    lib().lbm_release_cached()


def last_timing() -> dict:
    t = LbmTiming()
    lib().lbm_get_timing(C.byref(t))
    out = {"total_ms": t.total_ms, "kernels": {}}
    for k in range(t.n_kernels):
        out["kernels"][t.names[k].value.decode()] = {"avg_ms": t.avg_ms[k], "launches": int(t.launches[k])}
    return out


class Solver:
    """One solve on the GPU, built from a uq_amd.cases.Case."""

    def __init__(self, case, precision: str = "fp64", init: bool = True):
        self.case = case
        self.deck = PackedDeck(case.params, case.bnd, case.props)
        self.nx, self.ny = case.nx, case.ny
        h = C.c_void_p()
        _check(lib().lbm_create(C.byref(self.deck.params), self.nx, self.ny, self.deck.bnd_ptr(),
                                C.byref(self.deck.props), PRECISIONS[precision], C.byref(h)),
               "lbm_create")
        self._h = h
        self.precision = precision
        v = case.meta.get("viscosity") if isinstance(case.meta, dict) else None
        if v:
            self.set_viscosity(v["model"], v["mu_ref"], v["T_ref"], v.get("param", 0.0))
        sp = case.meta.get("spacing") if isinstance(case.meta, dict) else None
        if sp:
            self.set_spacing(sp["dx"], sp["dy"])
        if init:
            self.init()

    def set_spacing(self, dx, dy) -> None:
        """Non-uniform grid: dx per column (nx) and dy per row (ny), before
        the first step; decks with solid cells are refused."""
        dx = np.ascontiguousarray(dx, dtype=np.float64)
        dy = np.ascontiguousarray(dy, dtype=np.float64)
        if dx.shape != (self.nx,) or dy.shape != (self.ny,):
            raise ValueError("spacing arrays must have nx and ny entries")
        _check(lib().lbm_set_spacing(self._h, dptr(dx), dptr(dy)), "lbm_set_spacing")

    def set_viscosity(self, model, mu_ref: float, T_ref: float, param: float = 0.0) -> None:
        """Viscosity law: "constant", "vhs" (exponent 0.5 + param) or
        "sutherland" (constant S = param)."""
        code = {"constant": 0, "vhs": 1, "sutherland": 2}[model] if isinstance(model, str) else int(model)
        _check(lib().lbm_set_viscosity(self._h, code, float(mu_ref), float(T_ref), float(param)),
               "lbm_set_viscosity")

    @property
    def handle(self):
        return self._h

    def init(self):
        _check(lib().lbm_init(self._h), "lbm_init")

    def step(self, n: int, want_residual: bool = False):
        res = C.c_double(0.0)
        st = _check(lib().lbm_step(self._h, int(n), int(want_residual), C.byref(res)), "lbm_step",
                    allow_failed=True)
        return (st, res.value) if want_residual else st

    def info(self) -> dict:
        i = LbmInfo()
        _check(lib().lbm_get_info(self._h, C.byref(i)), "lbm_get_info")
        return {k: getattr(i, k) for k, _ in LbmInfo._fields_}

    def get_macro(self) -> dict:
        n = self.nx * self.ny
        arr = {k: np.empty(n) for k in ("rho", "ux", "uy", "p", "T")}
        _check(lib().lbm_get_macro(self._h, *(dptr(arr[k]) for k in ("rho", "ux", "uy", "p", "T"))),
               "lbm_get_macro")
        return {k: v.reshape(self.nx, self.ny) for k, v in arr.items()}

    def totals(self) -> np.ndarray:
        """[mass, x momentum, y momentum, total energy] over the fluid cells."""
        out = np.zeros(4)
        _check(lib().lbm_get_totals(self._h, dptr(out)), "lbm_get_totals")
        return out

    def get_dist(self):
        n = self.nx * self.ny * LBM_Q
        f, h = np.empty(n), np.empty(n)
        _check(lib().lbm_get_dist(self._h, dptr(f), dptr(h)), "lbm_get_dist")
        return f.reshape(self.nx, self.ny, LBM_Q), h.reshape(self.nx, self.ny, LBM_Q)

    def set_dist(self, f, h):
        f = np.ascontiguousarray(f, dtype=np.float64).reshape(-1)
        h = np.ascontiguousarray(h, dtype=np.float64).reshape(-1)
        _check(lib().lbm_set_dist(self._h, dptr(f), dptr(h)), "lbm_set_dist")

    def close(self):
        if self._h:
            lib().lbm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Several solvers advanced together, one launch per RK stage for all."""

    def __init__(self, solvers):
        self.solvers = list(solvers)
        arr = (C.c_void_p * len(self.solvers))(*[s.handle for s in self.solvers])
        h = C.c_void_p()
        _check(lib().lbm_batch_create(arr, len(self.solvers), C.byref(h)), "lbm_batch_create")
        self._h = h

    def step(self, n: int) -> int:
        return _check(lib().lbm_batch_step(self._h, int(n)), "lbm_batch_step", allow_failed=True)

    def autotune(self, steps: int = 5) -> tuple[int, list[float]]:
        """Pick the fastest flux tile shape for this batch (the lowest-numbered
        shape within 0.5 % of the fastest; state unchanged);
        returns the shape and each shape's milliseconds per step."""
        shape = C.c_int32(0)
        ms = np.zeros(FLUX_SHAPES)
        _check(lib().lbm_batch_autotune(self._h, int(steps), C.byref(shape), dptr(ms)), "lbm_batch_autotune")
        return shape.value, ms.tolist()

    def close(self):
        if self._h:
            lib().lbm_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching request loop (see lbm_server_* in uqlbm.h)."""

    def __init__(self, precision: str = "mixed", rk_method: int = 1, hbm_budget_bytes: int = 200 << 30,
                 max_slots: int = 1024, keep_results: bool = False):
        h = C.c_void_p()
        _check(lib().lbm_server_create(PRECISIONS[precision], rk_method, hbm_budget_bytes, max_slots,
                                       C.byref(h)), "lbm_server_create")
        self._h = h
        self._decks = {}
        if keep_results:
            _check(lib().lbm_server_keep_results(self._h, 1), "lbm_server_keep_results")

    def submit(self, case, tag: int, nsteps: int | None = None) -> None:
        deck = PackedDeck(case.params, case.bnd, case.props)
        n = case.nsteps if nsteps is None else nsteps
        _check(lib().lbm_server_submit(self._h, C.byref(deck.params), case.nx, case.ny, deck.bnd_ptr(),
                                       C.byref(deck.props), int(n), int(tag)), "lbm_server_submit")
        self._decks[tag] = (case.nx, case.ny)

    def tick(self, steps: int):
        a, q = C.c_int32(0), C.c_int32(0)
        _check(lib().lbm_server_tick(self._h, int(steps), C.byref(a), C.byref(q)), "lbm_server_tick",
               allow_failed=True)
        return a.value, q.value

    def poll(self, max_results: int = 4096) -> list[dict]:
        buf = (LbmRequestResult * max_results)()
        n = lib().lbm_server_poll(self._h, buf, max_results)
        if n < 0:
            raise LbmError(n, "lbm_server_poll")
        return [{k: getattr(buf[i], k) for k, _ in LbmRequestResult._fields_} for i in range(n)]

    def result_macro(self, tag: int) -> dict:
        nx, ny = self._decks[tag]
        arr = {k: np.empty(nx * ny) for k in ("rho", "ux", "uy", "p", "T")}
        _check(lib().lbm_server_result_macro(self._h, int(tag), *(dptr(arr[k]) for k in
                                                                  ("rho", "ux", "uy", "p", "T"))),
               "lbm_server_result_macro")
        return {k: v.reshape(nx, ny) for k, v in arr.items()}

    def stats(self) -> dict:
        lups, hbm = C.c_int64(0), C.c_size_t(0)
        _check(lib().lbm_server_stats(self._h, C.byref(lups), C.byref(hbm)), "lbm_server_stats")
        return {"lups": lups.value, "hbm_in_use": hbm.value}

    def close(self):
        if self._h:
            lib().lbm_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def clock_s() -> float:
    return float(lib().lbm_clock_s())


def rel_l2(a, b) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b.ravel())
    num = np.linalg.norm((a - b).ravel())
    return float(num / den) if den > 0 else float(num)
