"""HIP graph replay of the step sequence gives bitwise the results of direct
launches: plain solvers (RK3 and RK1), a batch of uneven instances, and
the server; each replayed several times with ordinary work (readbacks,
direct steps) in between."""
import numpy as np
import pytest

from uq_amd import cases

pytestmark = pytest.mark.gpu


def _run_solver(lbm, case, chunks, graphs, precision="fp64"):
    lbm.set_graphs(graphs)
    try:
        s = lbm.Solver(case, precision)
        snaps = []
        for n in chunks:
            assert s.step(n) == 0
            f, h = s.get_dist()
            snaps.append((f, h, s.info()["step"]))
        s.close()
        return snaps
    finally:
        lbm.set_graphs(True)


@pytest.mark.parametrize("case", [cases.couette(40, steps=50), cases.shear_wave(32, rk_method=0),
                                  cases.flat_plate(48, 32, plate_start=12, steps=20)],
                         ids=lambda c: f"{c.name}_rk{c.params['rk_method']}")
def test_solver_graph_replays_equal_direct(gpu_lib, case):
    chunks = [7, 1, 4, 3]
    print(f"[graphs] solver {case.name}", flush=True)
    a = _run_solver(gpu_lib, case, chunks, graphs=True)
    b = _run_solver(gpu_lib, case, chunks, graphs=False)
    for (fa, ha, sa), (fb, hb, sb) in zip(a, b):
        assert sa == sb
        assert np.array_equal(fa, fb) and np.array_equal(ha, hb)


def test_batch_graph_replays_equal_direct(gpu_lib):
    members = cases.uq_ensemble(6, seed=11, sides=(64, 128), steps=50)
    print("[graphs] batch of 6", flush=True)
    outs = {}
    for graphs in (True, False):
        gpu_lib.set_graphs(graphs)
        try:
            ss = [gpu_lib.Solver(c, "mixed") for c in members]
            b = gpu_lib.Batch(ss)
            for n in (5, 2, 3):
                assert b.step(n) == 0
                _ = ss[0].get_macro()   # direct work between replays
            outs[graphs] = [s.get_dist() for s in ss]
            b.close()
            for s in ss:
                s.close()
        finally:
            gpu_lib.set_graphs(True)
    for (fa, ha), (fb, hb) in zip(outs[True], outs[False]):
        assert np.array_equal(fa, fb) and np.array_equal(ha, hb)


def test_server_graph_replays_equal_direct(gpu_lib):
    reqs = [c.with_steps(n) for c, n in zip(cases.uq_ensemble(4, seed=3, sides=(64,), steps=30), (30, 12, 21, 7))]
    print("[graphs] server of 4", flush=True)
    res = {}
    for graphs in (True, False):
        gpu_lib.set_graphs(graphs)
        try:
            sv = gpu_lib.Server("fp64", max_slots=8, keep_results=True)
            for tag, c in enumerate(reqs):
                sv.submit(c, tag)
            done = []
            while len(done) < len(reqs):
                sv.tick(5)
                done += sv.poll()
            res[graphs] = {d["tag"]: sv.result_macro(d["tag"]) for d in done}
            sv.close()
        finally:
            gpu_lib.set_graphs(True)
    for tag in res[True]:
        for k in ("rho", "ux", "uy", "T"):
            assert np.array_equal(res[True][tag][k], res[False][tag][k]), (tag, k)
