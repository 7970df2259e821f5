"""Batched stepping and the continuous-batching server on the GPU.

Per-instance arithmetic does not depend on the batch: every member of a
batch, and every request served in a mix, must equal the same solve run
alone bit for bit — also when the group runs its flux stages on the strip
kernels and the lone run on the tile kernels (lbm_set_strips 2 and 0; the
default rule gives these small groups tiles), a group that mixes strip
instances with tile instances (the plate needs the physical path).
"""
import numpy as np
import pytest

from uq_amd import cases

pytestmark = pytest.mark.gpu


def _small_ensemble():
    rng = np.random.default_rng(11)
    out = []
    for k, (kind, n) in enumerate([("couette", 40), ("shear", 48), ("couette", 72), ("shear", 33),
                                   ("couette", 64), ("shear", 20)]):
        s = cases.uq_sample(rng)
        out.append(cases.uq_instance(kind, n, s, steps=30))
    out.append(cases.flat_plate(nx=70, ny=34, plate_start=18, steps=30))
    out.append(cases.riemann(50, steps=30))
    return out


# (precision, strip mode of the batch, strip mode of the lone runs): the
# mixed ensemble holds a shock deck whose strip stencils are FP64 and tile
# stencils FP32, so its cross-kernel case compares strips with strips
@pytest.mark.parametrize("precision,batch_mode,alone_mode",
                         [("fp64", 1, 1), ("mixed", 1, 1), ("fp64", 2, 0), ("mixed", 2, 2)],
                         ids=["fp64", "mixed", "fp64_strips_vs_tiles", "mixed_strips"])
def test_batch_equals_alone_bitwise(gpu_lib, precision, batch_mode, alone_mode):
    lbm = gpu_lib
    ens = _small_ensemble()
    print(f"[batch] {len(ens)} members precision={precision} modes {batch_mode}/{alone_mode}", flush=True)
    try:
        lbm.set_strips(alone_mode)
        alone = []
        for c in ens:
            s = lbm.Solver(c, precision)
            assert s.step(30) == 0
            alone.append(s.get_dist())
            s.close()
        lbm.set_strips(batch_mode)
        solvers = [lbm.Solver(c, precision) for c in ens]
        b = lbm.Batch(solvers)
        assert b.step(12) == 0
        assert b.step(18) == 0
        for s, (fa, ha) in zip(solvers, alone):
            f, h = s.get_dist()
            assert s.info()["step"] == 30
            assert np.array_equal(f, fa) and np.array_equal(h, ha)
    finally:
        lbm.set_strips(1)


def test_batch_member_failure_is_isolated(gpu_lib):
    lbm = gpu_lib
    good = cases.couette(40, steps=10)
    bad = cases.couette(40, steps=10)
    bad.props = {k: v.copy() for k, v in bad.props.items()}
    bad.props["density"][:] = 1e11       # equilibrium mass check trips
    bad.props["pressure"][:] = 0.0
    s_good, s_bad = lbm.Solver(good), lbm.Solver(bad)
    ref = lbm.Solver(good)
    ref.step(10)
    b = lbm.Batch([s_good, s_bad])
    st = b.step(10)
    assert st == 1                       # LBM_FAILED
    assert s_bad.info()["step"] == 1     # stopped at the failing step
    assert s_good.info()["step"] == 10
    fg, hg = s_good.get_dist()
    fr, hr = ref.get_dist()
    assert np.array_equal(fg, fr) and np.array_equal(hg, hr)
    # a failed solver refuses further steps
    assert s_bad.step(1) == 1


@pytest.mark.parametrize("server_mode,alone_mode", [(1, 1), (2, 0)], ids=["default", "strips_vs_tiles"])
def test_server_results_equal_alone(gpu_lib, server_mode, alone_mode):
    lbm = gpu_lib
    lbm.set_strips(server_mode)
    try:
        _serve_and_compare(lbm, alone_mode)
    finally:
        lbm.set_strips(1)


def _serve_and_compare(lbm, alone_mode):
    rng = np.random.default_rng(5)
    reqs = []
    for k in range(7):
        n = int(rng.integers(24, 90))
        steps = int(rng.integers(5, 40))
        kind = "couette" if k % 2 == 0 else "shear"
        reqs.append(cases.uq_instance(kind, n, cases.uq_sample(rng), steps=steps))
    sv = lbm.Server(precision="mixed", hbm_budget_bytes=1 << 34, max_slots=3, keep_results=True)
    for k, c in enumerate(reqs[:4]):
        sv.submit(c, tag=k)
    results = []
    active, queued = sv.tick(7)
    results += sv.poll()
    for k, c in enumerate(reqs[4:], start=4):   # arrivals while others are on the GPU
        sv.submit(c, tag=k)
    for _ in range(200):
        active, queued = sv.tick(9)
        results += sv.poll()
        assert active <= 3
        if not active and not queued:
            break
    assert sorted(r["tag"] for r in results) == list(range(len(reqs)))
    lbm.set_strips(alone_mode)
    for r in results:
        c = reqs[r["tag"]]
        assert r["status"] == 0 and r["steps_done"] == c.nsteps
        assert r["finish_s"] >= r["admit_s"] >= r["submit_s"]
        s = lbm.Solver(c, "mixed")
        s.step(c.nsteps)
        m = s.get_macro()
        mm = sv.result_macro(r["tag"])
        for k in ("rho", "ux", "uy", "T", "p"):
            assert np.array_equal(m[k], mm[k]), (r["tag"], k)
        s.close()
    assert sv.stats()["hbm_in_use"] == 0


def test_server_admits_by_memory_budget(gpu_lib):
    lbm = gpu_lib
    c = cases.couette(64, steps=6)
    from uq_amd.capi import LBM_PREC_MIXED  # noqa: F401
    one = lbm.Solver(c, "mixed")
    fp = one.info()["device_bytes"]
    one.close()
    sv = lbm.Server(precision="mixed", hbm_budget_bytes=int(2.5 * fp), max_slots=64)
    for k in range(6):
        sv.submit(c, tag=k)
    seen = 0
    done = []
    for _ in range(50):
        active, queued = sv.tick(3)
        seen = max(seen, active)
        done += sv.poll()
        if not active and not queued:
            break
    assert seen == 2
    assert len(done) == 6 and all(r["status"] == 0 for r in done)


def _normalised_residual_alone(lbm, case, steps):
    """Solve `case` alone for `steps` steps and apply the reference driver's
    normalisation (main.cpp:54-110): the residual of the last step over the
    largest residual of steps 2..20."""
    s = lbm.Solver(case, "mixed")
    hist = []
    for k in range(1, 21):
        _, r = s.step(1, want_residual=True)
        if k >= 2:
            hist.append(r)
    _, r = s.step(steps - 20, want_residual=True)
    m = s.get_macro()
    s.close()
    return r / max(hist), m


def test_server_retires_converged_request_early(gpu_lib):
    """A request whose deck carries a residual tolerance retires at the first
    round end where the normalised residual is <= tol, well before its step
    count; its fields equal the same solve run alone for the same number of
    steps, and the residual it reports is the one a lone solve computes.
    A run-to-tolerance request (step count 0) and a plain request share the
    server with it."""
    lbm = gpu_lib
    conv = cases.couette(24, nx=16, u_wall=0.1, tau=0.05, steps=4000)
    conv.params["Tc_Tref"] = 2.5
    conv.params["tol"] = 2e-2
    open_ended = cases.couette(20, nx=8, u_wall=0.08, tau=0.05, steps=0)
    open_ended.params["Tc_Tref"] = 2.5
    open_ended.params["tol"] = 5e-2
    plain = cases.shear_wave(32, u0=0.1, steps=300)
    sv = lbm.Server(precision="mixed", hbm_budget_bytes=1 << 34, max_slots=8, keep_results=True)
    try:
        sv.submit(conv, tag=0, nsteps=4000)
        sv.submit(open_ended, tag=1, nsteps=0)
        sv.submit(plain, tag=2)
        out = []
        for _ in range(400):
            active, queued = sv.tick(50)
            out += sv.poll()
            if not active and not queued:
                break
        by = {r["tag"]: r for r in out}
        assert sorted(by) == [0, 1, 2]
        assert by[2]["status"] == 0 and by[2]["steps_done"] == 300 and by[2]["residual"] == 0.0
        for tag, case in ((0, conv), (1, open_ended)):
            r = by[tag]
            tol = case.params["tol"]
            print(f"[serve] request {tag}: retired after {r['steps_done']} steps, normalised residual "
                  f"{r['residual']:.4e} (tol {tol})", flush=True)
            assert r["status"] == 0 and 21 <= r["steps_done"] < 4000
            assert r["residual"] <= tol
            res_alone, m_alone = _normalised_residual_alone(lbm, case, r["steps_done"])
            assert res_alone == pytest.approx(r["residual"], rel=1e-9)
            m = sv.result_macro(tag)
            for k in ("rho", "ux", "uy", "T", "p"):
                assert np.array_equal(m[k], m_alone[k]), (tag, k)
    finally:
        sv.close()


def test_destroying_a_batched_solver_is_deferred(gpu_lib):
    lbm = gpu_lib
    a, b = lbm.Solver(cases.couette(32, steps=10)), lbm.Solver(cases.shear_wave(24, steps=10))
    ref = lbm.Solver(cases.shear_wave(24, steps=10))
    batch = lbm.Batch([a, b])
    a.close()                      # freed only when the batch goes
    assert batch.step(10) == 0
    ref.step(10)
    fb, hb = b.get_dist()
    fr, hr = ref.get_dist()
    assert np.array_equal(fb, fr) and np.array_equal(hb, hr)
    batch.close()
    b.close()
    ref.close()
