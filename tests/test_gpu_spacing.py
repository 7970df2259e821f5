"""Non-uniform grids (lbm_set_spacing; the OpenCL front end's per-column dx
and per-row dy, PY/codeCL.py:611-618, 676-679) on the GPU: uniform arrays
reproduce the uniform path bit for bit; stretched grids match the oracle
within max(1e-12, 8 x the oracle's one-ulp sensitivity) relative L2 (a
doubly periodic deck whose state varies along both axes, and the clamped
Riemann deck, and decks with walls: the C++ wall closure with the wall cell's
own dx, dy); a batch mixing spaced and uniform members gives each member's
result alone; non-positive spacings and late calls are refused."""
import numpy as np
import pytest

from uq_amd import cases
from uq_amd.lbm import LbmError, rel_l2

pytestmark = pytest.mark.gpu


def _modulate(get_dist, set_dist, nx, ny):
    f, h = get_dist()
    m = 1 + 0.02 * np.outer(np.sin(2 * np.pi * np.arange(nx) / nx), np.cos(2 * np.pi * np.arange(ny) / ny))
    set_dist(f * m[:, :, None], h * m[:, :, None])


def _gpu(lbm, case, steps, precision="fp64", modulate=False):
    s = lbm.Solver(case, precision)
    try:
        if modulate:
            _modulate(s.get_dist, s.set_dist, case.nx, case.ny)
        st = s.step(steps)
        f, h = s.get_dist()
        return st, f, h
    finally:
        s.close()


def _oracle(case, steps, modulate=False, perturb=False):
    from pyoracle import Oracle
    o = Oracle(case)
    try:
        if modulate:
            _modulate(o.get_dist, o.set_dist, case.nx, case.ny)
        if perturb:  # one ulp, random signs: the oracle's own sensitivity
            f, h = o.get_dist()
            rng = np.random.default_rng(1)
            eps = np.finfo(np.float64).eps
            o.set_dist(f * (1 + eps * rng.choice([-1.0, 1.0], f.shape)),
                       h * (1 + eps * rng.choice([-1.0, 1.0], h.shape)))
        st = o.step(steps)
        f, h = o.get_dist()
        return st, f, h
    finally:
        o.close()


def test_uniform_arrays_are_the_uniform_path(gpu_lib):
    c = cases.shear_wave(40, steps=12)
    ref = _gpu(gpu_lib, c, 12)
    got = _gpu(gpu_lib, cases.with_spacing(c, np.full(c.nx, c.params["dx"]), np.full(c.ny, c.params["dy"])), 12)
    assert ref[0] == 0 and got[0] == 0
    assert np.array_equal(ref[1], got[1]) and np.array_equal(ref[2], got[2])


@pytest.mark.parametrize("deck", ["periodic", "riemann"])
def test_stretched_grid_matches_oracle(gpu_lib, deck):
    if deck == "periodic":
        c, steps, mod = cases.stretched(cases.shear_wave(48, u0=0.15, steps=30), amp=0.3), 30, True
    else:
        c, steps, mod = cases.stretched(cases.riemann(64, steps=6), amp=0.2), 6, False
    st_o, fo, ho = _oracle(c, steps, mod)
    st_g, fg, hg = _gpu(gpu_lib, c, steps, modulate=mod)
    ef, eh = rel_l2(fg, fo), rel_l2(hg, ho)
    # bar: max(1e-12, 8 x the oracle's one-ulp sensitivity), as for the
    # golden fixtures (the Riemann shocks amplify rounding)
    _, fp, hp = _oracle(c, steps, mod, perturb=True)
    bar = max(1e-12, 8.0 * max(rel_l2(fp, fo), rel_l2(hp, ho)))
    print(f"[spacing] {c.name} {steps} steps: f {ef:.2e} h {eh:.2e} bar {bar:.2e}", flush=True)
    assert st_o == 0 and st_g == 0
    assert ef <= bar and eh <= bar
    # the spacing matters: the uniform deck ends elsewhere
    st_u, fu, _ = _gpu(gpu_lib, cases.shear_wave(48, u0=0.15, steps=30) if deck == "periodic"
                       else cases.riemann(64, steps=6), steps, modulate=mod)
    assert rel_l2(fu, fo) > 1e-6


@pytest.mark.parametrize("rk,fm", [(0, 1), (2, 1), (3, 1), (1, 0)], ids=["rk1", "euler", "euler_eq", "rk3_nnd"])
def test_stretched_grid_other_steppers_and_nnd(gpu_lib, rk, fm):
    """The spacing enters the net flux that every stepper and both
    reconstructions share (RKMETHOD 0/2/3, FMETHOD 0)."""
    c = cases.stretched(cases.shear_wave(40, u0=0.12, steps=20, rk_method=rk, flux_method=fm), amp=0.3)
    st_o, fo, ho = _oracle(c, 20, True)
    st_g, fg, hg = _gpu(gpu_lib, c, 20, modulate=True)
    _, fp, hp = _oracle(c, 20, True, perturb=True)
    bar = max(1e-12, 8.0 * max(rel_l2(fp, fo), rel_l2(hp, ho)))
    ef, eh = rel_l2(fg, fo), rel_l2(hg, ho)
    print(f"[spacing] {c.name} rk{rk} fm{fm}: f {ef:.2e} h {eh:.2e} bar {bar:.2e}", flush=True)
    assert st_o == 0 and st_g == 0
    assert ef <= bar and eh <= bar


def test_batch_mixing_spaced_and_uniform_members(gpu_lib):
    a = cases.stretched(cases.shear_wave(32, u0=0.12, steps=20), amp=0.25)
    b = cases.shear_wave(48, u0=0.1, steps=20)
    alone = [_gpu(gpu_lib, c, 20) for c in (a, b)]
    ss = [gpu_lib.Solver(c, "fp64") for c in (a, b)]
    bt = gpu_lib.Batch(ss)
    try:
        assert bt.step(20) == 0
        for s, (st, f, h) in zip(ss, alone):
            fb, hb = s.get_dist()
            assert np.array_equal(fb, f) and np.array_equal(hb, h)
    finally:
        bt.close()
        for s in ss:
            s.close()


def test_spacing_set_after_init_on_a_batch_member(gpu_lib):
    """Spacing set after initialisation, on a solver already in a batch but
    before its first step, reaches the batched kernels (the batch reads its
    members' instance records at every step call)."""
    c = cases.shear_wave(32, u0=0.12, steps=20)
    sp = cases.stretched(c, amp=0.25)
    st, f_ref, h_ref = _gpu(gpu_lib, sp, 20)
    s = gpu_lib.Solver(c, "fp64")
    bt = gpu_lib.Batch([s])
    try:
        s.set_spacing(sp.meta["spacing"]["dx"], sp.meta["spacing"]["dy"])
        assert st == 0 and bt.step(20) == 0
        f, h = s.get_dist()
        assert np.array_equal(f, f_ref) and np.array_equal(h, h_ref)
    finally:
        bt.close()
        s.close()


def test_mixed_precision_on_a_stretched_grid(gpu_lib):
    """Mixed precision's bar is on the flow fields (rho, u, T) against the
    FP64 path, not on the distributions, whose relative L2 is dominated by
    the rest-state populations."""
    c = cases.stretched(cases.shear_wave(48, u0=0.15, steps=30), amp=0.3)
    out = {}
    for prec in ("fp64", "mixed"):
        s = gpu_lib.Solver(c, prec)
        try:
            assert s.step(30) == 0
            out[prec] = s.get_macro()
        finally:
            s.close()
    m32, m64 = out["mixed"], out["fp64"]
    e = {k: rel_l2(m32[k], m64[k]) for k in ("rho", "T")}
    e["u"] = rel_l2(np.stack([m32["ux"], m32["uy"]]), np.stack([m64["ux"], m64["uy"]]))
    print(f"[spacing] mixed vs fp64 over 30 steps: {e}", flush=True)
    assert all(v <= 1e-6 for v in e.values()), e


@pytest.mark.parametrize("deck", ["obstacle", "couette"])
def test_stretched_grid_with_walls_matches_oracle(gpu_lib, deck):
    """Walls on a non-uniform grid: flow past a block (its corner cells
    have fluid on both axes, where the wall balance's x and y weights do not
    cancel) and the Couette channel, against the oracle; the area-weighted
    fluid mass is conserved (the balance's purpose)."""
    if deck == "obstacle":
        base = cases.obstacle(40, block=(15, 15, 6, 6), u0=0.15, steps=30)
    else:
        base = cases.couette(40, u_wall=0.2, steps=30)
    c = cases.stretched(base, amp=0.3)
    st_o, fo, ho = _oracle(c, 30, True)
    _, fp, hp = _oracle(c, 30, True, perturb=True)
    bar = max(1e-12, 8.0 * max(rel_l2(fp, fo), rel_l2(hp, ho)))
    s = gpu_lib.Solver(c, "fp64")
    try:
        _modulate(s.get_dist, s.set_dist, c.nx, c.ny)
        f0, _ = s.get_dist()
        st_g = s.step(30)
        fg, hg = s.get_dist()
    finally:
        s.close()
    ef, eh = rel_l2(fg, fo), rel_l2(hg, ho)
    print(f"[spacing] {c.name} with walls: f {ef:.2e} h {eh:.2e} bar {bar:.2e}", flush=True)
    assert st_o == 0 and st_g == 0
    assert ef <= bar and eh <= bar
    vol = np.outer(c.meta["spacing"]["dx"], c.meta["spacing"]["dy"])
    fluid = np.asarray(c.bnd) == 0 if deck == "obstacle" else np.pad(np.ones((c.nx, c.ny - 2), bool), ((0, 0), (1, 1)))
    m0, m1 = (vol * f0.sum(axis=2))[fluid].sum(), (vol * fg.sum(axis=2))[fluid].sum()
    print(f"[spacing] {c.name}: weighted fluid mass drift {abs(m1 - m0) / m0:.1e}", flush=True)
    assert abs(m1 - m0) <= 1e-13 * m0


def test_refusals(gpu_lib):
    c = cases.shear_wave(16, steps=5)
    s = gpu_lib.Solver(c, "fp64")
    try:
        with pytest.raises(LbmError):
            s.set_spacing(np.full(c.nx, c.params["dx"]), -np.ones(c.ny))
        s.step(1)
        with pytest.raises(LbmError):
            s.set_spacing(np.full(c.nx, c.params["dx"]), np.full(c.ny, c.params["dy"]))
    finally:
        s.close()
