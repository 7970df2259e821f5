"""ctypes wrapper of the CPU oracle (liblbm_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke()
and by bench.py's cpu_baseline leg, as the checker and the reported CPU
baseline. The product package never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from uq_amd.capi import LBM_Q, PackedDeck, dptr

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblbm_oracle.so")

_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc -O2, no fast-math)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE, "-B" if force else "all"], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run `make -C oracle`)")
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        d, i32, vp = C.c_double, C.c_int32, C.c_void_p
        L.orc_create.restype = vp
        L.orc_create.argtypes = [vp, i32, i32, P(i32), vp]
        L.orc_destroy.argtypes = [vp]
        L.orc_step.restype = C.c_int
        L.orc_step.argtypes = [vp, i32, i32, P(d)]
        L.orc_get_dist.argtypes = [vp, P(d), P(d)]
        L.orc_set_dist.argtypes = [vp, P(d), P(d)]
        L.orc_get_macro.argtypes = [vp, P(d), P(d), P(d), P(d), P(d)]
        L.orc_get_consts.argtypes = [vp, P(d)]
        L.orc_equilibrium.argtypes = [vp, d, d, d, d, P(d), P(d), P(i32)]
        L.orc_weno5.restype = d
        L.orc_weno5.argtypes = [d, P(d)]
        L.orc_nnd.restype = d
        L.orc_nnd.argtypes = [d, P(d)]
        L.orc_minmod.restype = d
        L.orc_minmod.argtypes = [d, d]
        L.orc_index.restype = i32
        L.orc_index.argtypes = [vp, i32, i32, i32, P(i32)]
        L.orc_mirror_dirs.argtypes = [i32, i32, i32, P(i32), P(i32)]
        L.orc_stencil.argtypes = [vp, P(d), i32, i32, i32, i32, P(d), P(d)]
        L.orc_boundary_fluxes.argtypes = [vp, P(d), P(d), P(d), P(d)]
        L.orc_cflux.restype = d
        L.orc_cflux.argtypes = [vp, P(d), i32, i32, i32]
        L.orc_step_seconds.restype = d
        L.orc_step_seconds.argtypes = [vp]
        L.orc_set_viscosity.argtypes = [vp, i32, d, d, d]
        L.orc_set_spacing.argtypes = [vp, P(d), P(d)]
        _lib = L
    return _lib


def weno5(ee: float, S) -> float:
    S = np.ascontiguousarray(S, dtype=np.float64)
    return lib().orc_weno5(ee, dptr(S))


def nnd(ee: float, S) -> float:
    S = np.ascontiguousarray(S, dtype=np.float64)
    return lib().orc_nnd(ee, dptr(S))


def minmod(a: float, b: float) -> float:
    return lib().orc_minmod(a, b)


def mirror_dirs(i: int, mir_x: int, mir_y: int):
    a, b = C.c_int32(), C.c_int32()
    lib().orc_mirror_dirs(i, mir_x, mir_y, C.byref(a), C.byref(b))
    return a.value, b.value


class Oracle:
    """One oracle solver instance built from a Case (uq_amd.cases)."""

    def __init__(self, case):
        self.case = case
        self.deck = PackedDeck(case.params, case.bnd, case.props)
        self.nx, self.ny = case.nx, case.ny
        self._h = lib().orc_create(C.byref(self.deck.params), self.nx, self.ny,
                                   self.deck.bnd_ptr(), C.byref(self.deck.props))
        if not self._h:
            raise ValueError("orc_create rejected the deck")
        v = case.meta.get("viscosity") if isinstance(case.meta, dict) else None
        if v:
            self.set_viscosity(v["model"], v["mu_ref"], v["T_ref"], v.get("param", 0.0))
        sp = case.meta.get("spacing") if isinstance(case.meta, dict) else None
        if sp:
            self.set_spacing(sp["dx"], sp["dy"])

    def set_viscosity(self, model, mu_ref: float, T_ref: float, param: float = 0.0) -> None:
        """Viscosity law: "constant", "vhs" (exponent 0.5 + param) or
        "sutherland" (constant S = param)."""
        code = {"constant": 0, "vhs": 1, "sutherland": 2}[model] if isinstance(model, str) else int(model)
        if lib().orc_set_viscosity(self._h, code, float(mu_ref), float(T_ref), float(param)) != 0:
            raise ValueError("bad viscosity law")

    def set_spacing(self, dx, dy) -> None:
        """Per-column dx[nx] and per-row dy[ny] (non-uniform grid of the
        OpenCL front end); before the first step; no solid cells."""
        self._dx = np.ascontiguousarray(dx, dtype=np.float64)
        self._dy = np.ascontiguousarray(dy, dtype=np.float64)
        if self._dx.shape != (self.nx,) or self._dy.shape != (self.ny,):
            raise ValueError("spacing arrays must have nx and ny entries")
        if lib().orc_set_spacing(self._h, dptr(self._dx), dptr(self._dy)) != 0:
            raise ValueError("bad spacing (non-positive, or a deck with solid cells)")

    def close(self):
        if self._h:
            lib().orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n: int, want_res: bool = False):
        res = C.c_double(0.0)
        st = lib().orc_step(self._h, int(n), int(want_res), C.byref(res))
        return (st, res.value) if want_res else st

    def consts(self) -> dict:
        out = np.zeros(10)
        lib().orc_get_consts(self._h, dptr(out))
        keys = ("dt", "c", "Tc", "b", "K", "dx", "dy", "nsteps", "step", "tstep")
        return dict(zip(keys, out.tolist()))

    def get_dist(self):
        n = self.nx * self.ny * LBM_Q
        f, h = np.empty(n), np.empty(n)
        lib().orc_get_dist(self._h, dptr(f), dptr(h))
        return f.reshape(self.nx, self.ny, LBM_Q), h.reshape(self.nx, self.ny, LBM_Q)

    def set_dist(self, f, h):
        f = np.ascontiguousarray(f, dtype=np.float64).reshape(-1)
        h = np.ascontiguousarray(h, dtype=np.float64).reshape(-1)
        lib().orc_set_dist(self._h, dptr(f), dptr(h))

    def get_macro(self) -> dict:
        n = self.nx * self.ny
        arr = {k: np.empty(n) for k in ("rho", "ux", "uy", "p", "T")}
        lib().orc_get_macro(self._h, *(dptr(arr[k]) for k in ("rho", "ux", "uy", "p", "T")))
        return {k: v.reshape(self.nx, self.ny) for k, v in arr.items()}

    def equilibrium(self, rho, ux, uy, T):
        eqf, eqh = np.empty(LBM_Q), np.empty(LBM_Q)
        failed = C.c_int32(0)
        lib().orc_equilibrium(self._h, rho, ux, uy, T, dptr(eqf), dptr(eqh), C.byref(failed))
        return eqf, eqh, failed.value

    def index(self, i: int, pm: int, d: int):
        mir = C.c_int32(0)
        r = lib().orc_index(self._h, i, pm, d, C.byref(mir))
        return r, mir.value

    def stencil(self, f, x: int, y: int, i: int, length: int):
        f = np.ascontiguousarray(f, dtype=np.float64).reshape(-1)
        Sx, Sy = np.zeros(length), np.zeros(length)
        lib().orc_stencil(self._h, dptr(f), x, y, i, length, dptr(Sx), dptr(Sy))
        return Sx, Sy

    def boundary_fluxes(self, f, h):
        f = np.ascontiguousarray(f, dtype=np.float64).reshape(-1)
        h = np.ascontiguousarray(h, dtype=np.float64).reshape(-1)
        ff = np.zeros(2 * f.size)
        hf = np.zeros(2 * h.size)
        lib().orc_boundary_fluxes(self._h, dptr(f), dptr(h), dptr(ff), dptr(hf))
        shape = (self.nx, self.ny, LBM_Q, 2)
        return ff.reshape(shape), hf.reshape(shape)

    def cflux(self, flux, x: int, y: int, i: int) -> float:
        flux = np.ascontiguousarray(flux, dtype=np.float64).reshape(-1)
        return lib().orc_cflux(self._h, dptr(flux), x, y, i)

    def step_seconds(self) -> float:
        return lib().orc_step_seconds(self._h)
