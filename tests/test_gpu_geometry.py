"""Geometries where the tile kernels' logical stencil addressing differs from
the reference's physical addressing (cpuLBM.cpp:1171-1326 stencil and
extrapolation, 1436-1470 upwind owner): solid plates one cell thick and
solids one or two cells from a mirror edge. Their tiles are flagged and
their net fluxes formed with the reference's addressing (k_phys).

Bar (FP64): relative L2 of f, h, rho, u, T against the CPU oracle <=
max(1e-12, 8 x the oracle's own one-ulp sensitivity over the same steps).
Mixed precision: <= 1e-6 against the FP64 path.
"""
import numpy as np
import pytest

from uq_amd import cases
from uq_amd.lbm import rel_l2

pytestmark = pytest.mark.gpu

DECKS = {
    # a plate one cell thick across the flow, and one along it
    "plate_across": dict(block=(18, 20, 1, 8)),
    "plate_along": dict(block=(18, 20, 8, 1)),
    # solids one and two rows above a mirrored south edge
    "mirror_gap1": dict(block=(18, 1, 8, 3), mirror_south=True),
    "mirror_gap2_thin": dict(block=(18, 2, 8, 1), mirror_south=True),
}


def _deck(name, **kw):
    return cases.obstacle(48, u0=0.15, **DECKS[name], **kw)


def _oracle(case, steps, perturb=False):
    from pyoracle import Oracle
    o = Oracle(case)
    try:
        if perturb:
            f, h = o.get_dist()
            rng = np.random.default_rng(1)
            eps = np.finfo(np.float64).eps
            o.set_dist(f * (1 + eps * rng.choice([-1.0, 1.0], f.shape)),
                       h * (1 + eps * rng.choice([-1.0, 1.0], h.shape)))
        st = o.step(steps)
        f, h = o.get_dist()
        m = o.get_macro()
        return st, f, h, m
    finally:
        o.close()


def _errors(f, h, m, fo, ho, mo):
    return {"f": rel_l2(f, fo), "h": rel_l2(h, ho), "rho": rel_l2(m["rho"], mo["rho"]),
            "T": rel_l2(m["T"], mo["T"]),
            "u": rel_l2(np.stack([m["ux"], m["uy"]]), np.stack([mo["ux"], mo["uy"]]))}


def _check(gpu_lib, case, steps):
    so, fo, ho, mo = _oracle(case, steps)
    _, fp, hp, _ = _oracle(case, steps, perturb=True)
    bar = max(1e-12, 8.0 * max(rel_l2(fp, fo), rel_l2(hp, ho)))
    s = gpu_lib.Solver(case, "fp64")
    try:
        sg = s.step(steps)
        f, h = s.get_dist()
        m = s.get_macro()
    finally:
        s.close()
    e = _errors(f, h, m, fo, ho, mo)
    print(f"[geometry] {case.name} rk{case.params['rk_method']} fm{case.params['flux_method']} {steps} steps: "
          + " ".join(f"{k}={v:.2e}" for k, v in e.items()) + f" bar {bar:.2e}", flush=True)
    assert so == 0 and sg == 0
    assert all(v <= bar for v in e.values()), e
    return f, h


@pytest.mark.parametrize("name", list(DECKS))
def test_thin_and_mirror_edge_solids_match_oracle(gpu_lib, name):
    _check(gpu_lib, _deck(name, steps=30), 30)


@pytest.mark.parametrize("rk,fm", [(0, 1), (2, 1), (1, 0)], ids=["rk1", "euler", "rk3_nnd"])
def test_physical_path_other_steppers(gpu_lib, rk, fm):
    """Every flux launch (stage-2 with moment sums, stage-3, the one-stage
    steppers' single flux, NND) has its physical-addressing counterpart."""
    _check(gpu_lib, _deck("plate_along", steps=20, rk_method=rk, flux_method=fm), 20)
    _check(gpu_lib, _deck("mirror_gap1", steps=20, rk_method=rk, flux_method=fm), 20)


def test_thin_plate_in_a_batch_equals_alone(gpu_lib):
    """A batch mixing members with flagged tiles and members without gives
    each member its result alone."""
    members = [_deck("plate_across", steps=15), cases.couette(40, steps=15), _deck("mirror_gap2_thin", steps=15)]
    alone = []
    for c in members:
        s = gpu_lib.Solver(c, "fp64")
        s.step(15)
        alone.append(s.get_dist())
        s.close()
    solvers = [gpu_lib.Solver(c, "fp64") for c in members]
    b = gpu_lib.Batch(solvers)
    try:
        assert b.step(15) == 0
        for s, (fa, ha) in zip(solvers, alone):
            f, h = s.get_dist()
            assert np.array_equal(f, fa) and np.array_equal(h, ha)
    finally:
        b.close()
        for s in solvers:
            s.close()


def test_thin_plate_mixed_precision(gpu_lib):
    """FP32 storage through the physical path. The flow is started
    impulsively against the plate, which the scheme does not survive for
    long (the oracle's equilibrium check trips near step 85 even at
    u0 = 0.03): the comparison stays well inside that window."""
    c = cases.obstacle(48, block=(18, 20, 1, 8), u0=0.02, steps=40)
    out = {}
    for prec in ("fp64", "mixed"):
        s = gpu_lib.Solver(c, prec)
        try:
            assert s.step(40) == 0
            out[prec] = s.get_macro()
        finally:
            s.close()
    e = {k: rel_l2(out["mixed"][k], out["fp64"][k]) for k in ("rho", "T")}
    e["u"] = rel_l2(np.stack([out["mixed"]["ux"], out["mixed"]["uy"]]),
                    np.stack([out["fp64"]["ux"], out["fp64"]["uy"]]))
    print(f"[geometry] thin plate mixed vs fp64 over 40 steps: {e}", flush=True)
    assert all(v <= 1e-6 for v in e.values()), e
