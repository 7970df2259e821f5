"""MAT-file decks and snapshots in the reference's formats.

read_mat_deck mirrors cpuLBM::loadInput (reference CODE/C++/VS2010/simLB/
simLB/cpuLBM.cpp:46-281): it reads the deck variables by name, takes Nx and
Ny from the bndArray dimensions (columns, rows) and flips bndArray so that
y = 0 is the last MATLAB row (y_ = Ny-1-y, :254-266). write_mat_snapshot
mirrors cpuLBM::createMAT (:467-767): rho, ux, uy, p, T as Ny x Nx arrays
holding the [x][y] fields in MATLAB column order, t, dx, dy, and with
distributions=True also f, h (13 x Ny x Nx) and ex, ey (13 x 1).

The MAT I/O uses scipy.io (MATLAB v5 format); no MATLAB library is needed.
"""
from __future__ import annotations

import numpy as np
import scipy.io

from .cases import Case

_INT_KEYS = {
    "RKMETHOD": "rk_method", "FMETHOD": "flux_method", "steps": "steps",
    "periodicX": "periodic_x", "periodicY": "periodic_y", "mirrorNorth": "mirror_north",
    "mirrorSouth": "mirror_south", "mirrorEast": "mirror_east", "mirrorWest": "mirror_west",
}
_FLOAT_KEYS = {
    "CFL": "cfl", "dtau": "dtau", "tt_tref": "tt_tref", "tt_time": "tt_time", "tol": "tol",
    "gamma": "gamma", "R": "R", "mu": "mu", "Pr": "Pr", "dx": "dx", "dy": "dy", "Lx": "Lx",
    "Ly": "Ly", "rho_ref": "rho_ref", "p_ref": "p_ref", "T_ref": "T_ref", "Tc_Tref": "Tc_Tref",
}
_PROP_KEYS = {"density": "density", "pressure": "pressure", "temp": "temp",
              "velocityX": "velocity_x", "velocityY": "velocity_y", "cellType": "cell_type"}


def read_mat_deck(path: str, name: str | None = None) -> Case:
    """Read a reference MAT deck into a Case. Scalars absent from the deck
    are 0 (the reference leaves them uninitialised); the reference's own
    conversions apply: ints via (int) truncation of the stored double."""
    m = scipy.io.loadmat(path, squeeze_me=False)
    params = {}
    extra = {}
    for k, v in m.items():
        if k.startswith("__"):
            continue
        a = np.asarray(v, dtype=np.float64)
        if k in _INT_KEYS:
            params[_INT_KEYS[k]] = int(a.ravel()[0])
        elif k in _FLOAT_KEYS:
            params[_FLOAT_KEYS[k]] = float(a.ravel()[0])
        elif k not in _PROP_KEYS and k != "bndArray":
            extra[k] = a.ravel()[0] if a.size == 1 else a
    if "bndArray" not in m:
        raise ValueError(f"{path}: no bndArray")
    raw = np.asarray(m["bndArray"])
    ny, nx = raw.shape                       # dims[0] = Ny rows, dims[1] = Nx columns
    bnd = np.empty((nx, ny), dtype=np.int32)
    for y in range(ny):
        bnd[:, y] = raw[ny - 1 - y, :].astype(np.int32)
    nprops = int(np.asarray(m.get("numProps", [[0]])).ravel()[0]) if "numProps" in m else None
    props = {}
    for mk, pk in _PROP_KEYS.items():
        if mk not in m:
            raise ValueError(f"{path}: property column {mk} missing")
        col = np.asarray(m[mk], dtype=np.float64).ravel()
        props[pk] = col.astype(np.int32) if pk == "cell_type" else col
    n = len(props["cell_type"])
    if nprops is not None and nprops != n:
        raise ValueError(f"{path}: numProps = {nprops} but the property columns have {n} rows")
    from .cases import _params
    full = {k: 0 for k in _params()}
    full.update(params)
    meta = dict(kind="deck", source=path, extra=extra)
    steps = int(full.get("steps", 0))
    return Case(name or "deck", full, bnd, props, steps, meta)


def write_mat_snapshot(path: str, macro: dict, t: float, dx: float, dy: float, dist=None,
                       c: float | None = None) -> None:
    """Write a snapshot like cpuLBM::createMAT. `macro` holds [x][y] arrays
    rho, ux, uy, p, T; `dist` = (f, h) as [x][y][13] arrays; `c` the
    lattice speed (ex = c * lattice x, ey likewise) when dist is given."""
    nx, ny = macro["rho"].shape
    out = {}
    for k in ("rho", "ux", "uy", "p", "T"):
        # createMAT memcpy's the [x][y] buffer into an {Ny, Nx} array: MATLAB
        # element (r, c) is the value at x = c, y = r (no y flip on output)
        out[k] = np.ascontiguousarray(np.asarray(macro[k], dtype=np.float64).T)
    out["t"] = np.array([[float(t)]])
    out["dx"] = np.array([[float(dx)]])
    out["dy"] = np.array([[float(dy)]])
    if dist is not None:
        f, h = dist
        # {13, Ny, Nx}: element (i, y, x) = buffer[(x*Ny + y)*13 + i]
        out["f"] = np.ascontiguousarray(np.transpose(np.asarray(f, dtype=np.float64), (2, 1, 0)))
        out["h"] = np.ascontiguousarray(np.transpose(np.asarray(h, dtype=np.float64), (2, 1, 0)))
        lat_x = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1, 2, 0, -2, 0], dtype=np.float64)
        lat_y = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1, 0, 2, 0, -2], dtype=np.float64)
        cc = 1.0 if c is None else float(c)
        out["ex"] = (lat_x * cc).reshape(13, 1)
        out["ey"] = (lat_y * cc).reshape(13, 1)
    scipy.io.savemat(path, out)


def read_mat_snapshot(path: str) -> dict:
    """Read a snapshot written by write_mat_snapshot back into [x][y] arrays."""
    m = scipy.io.loadmat(path)
    out = {k: np.asarray(m[k]).T.copy() for k in ("rho", "ux", "uy", "p", "T")}
    for k in ("t", "dx", "dy"):
        out[k] = float(np.asarray(m[k]).ravel()[0])
    if "f" in m:
        out["f"] = np.transpose(np.asarray(m["f"]), (2, 1, 0)).copy()
        out["h"] = np.transpose(np.asarray(m["h"]), (2, 1, 0)).copy()
    return out


def write_bin_deck(path: str, case: Case, precision: int = 0) -> None:
    """Write the binary deck read by the C++ driver uq_amd/lib/lbm_run."""
    from .capi import PackedDeck
    import ctypes
    d = PackedDeck(case.params, case.bnd, case.props)
    with open(path, "wb") as fh:
        fh.write(b"UQLBMDK1")
        fh.write(bytes(ctypes.string_at(ctypes.addressof(d.params), ctypes.sizeof(d.params))))
        n = len(case.props["cell_type"])
        fh.write(np.array([case.nx, case.ny, n, precision], dtype=np.int32).tobytes())
        fh.write(np.ascontiguousarray(case.bnd, dtype=np.int32).tobytes())
        for k in ("density", "pressure", "temp", "velocity_x", "velocity_y"):
            fh.write(np.ascontiguousarray(case.props[k], dtype=np.float64).tobytes())
        fh.write(np.ascontiguousarray(case.props["cell_type"], dtype=np.int32).tobytes())


def read_bin_output(path: str) -> dict:
    """Read the output file of uq_amd/lib/lbm_run."""
    raw = open(path, "rb").read()
    if raw[:8] != b"UQLBMOUT":
        raise ValueError("not a UQLBMOUT file")
    nx, ny, step, stop = np.frombuffer(raw[8:24], dtype=np.int32)
    n = int(nx) * int(ny)
    body = np.frombuffer(raw[24:24 + 5 * 8 * n], dtype=np.float64).reshape(5, int(nx), int(ny))
    out = {k: body[i].copy() for i, k in enumerate(("rho", "ux", "uy", "p", "T"))}
    out.update(step=int(step), stop=int(stop))
    return out
