"""Full-size parity check of a benchmark deck — the four-quadrant Riemann
problem (4096^2) or the Mach-2 flat plate (512 x 256, NND, inflow column,
free-stream top row, wall from x = 64) — the GPU FP64 path against the CPU
oracle after a few steps, with the oracle's own sensitivity to a one-ulp
perturbation of the initial state measured over the same steps.

Not a pytest module: the oracle needs about 2.5 KB of host memory per cell
(41 GB at 4096^2, twice that with the perturbed run beside it) and about a
minute per step. Run from the repository root on the GPU machine:

    python tests/golden/parity_fullsize.py --deck riemann --n 4096 --steps 3
    python tests/golden/parity_fullsize.py --deck plate --steps 40

The two oracle runs go first, in two processes (the GPU is not touched
before they finish); then the GPU path runs the same deck. Bar per field:
relative L2 <= max(1e-12, 8 x the oracle's sensitivity), as for the golden
fixtures. Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import multiprocessing as mp
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, HERE)

from uq_amd import cases  # noqa: E402
import make_golden  # noqa: E402


def deck(name: str, n: int, steps: int):
    if name == "riemann":
        return cases.riemann(n, steps=steps)
    return cases.flat_plate(512, 256, steps=steps)   # the 512 x 256 Mach-2 plate deck, plate from x = 64


def _oracle(args):
    name, n, steps, perturb = args
    t = time.time()
    st, m = make_golden.run_oracle(deck(name, n, steps), steps, perturb=perturb)
    print(f"[fullsize] oracle {name} perturb={perturb} status={st} {time.time() - t:.0f} s", flush=True)
    return perturb, st, {k: m[k] for k in make_golden.FIELDS}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--deck", choices=("riemann", "plate"), default="riemann")
    ap.add_argument("--n", type=int, default=4096, help="Riemann side")
    ap.add_argument("--steps", type=int, default=3)
    a = ap.parse_args()
    t0 = time.time()
    with mp.get_context("fork").Pool(2) as pool:
        res = dict((p, (st, m)) for p, st, m in
                   pool.map(_oracle, [(a.deck, a.n, a.steps, False), (a.deck, a.n, a.steps, True)]))
    cpu_s = time.time() - t0

    from uq_amd import lbm  # the GPU is first touched here, after the oracle runs
    case = deck(a.deck, a.n, a.steps)
    s = lbm.Solver(case, "fp64")
    st_gpu = s.step(a.steps)
    m = s.get_macro()
    s.close()
    ref = res[False][1]
    err = make_golden.field_errors(m, ref)
    sens = make_golden.field_errors(res[True][1], ref)
    bar = {k: max(1e-12, 8.0 * sens[k]) for k in err}
    out = {"deck": case.name, "nx": case.nx, "ny": case.ny, "steps": a.steps,
           "oracle_status": res[False][0], "gpu_status": st_gpu,
           "rel_l2": {k: float(f"{v:.3e}") for k, v in err.items()},
           "oracle_1ulp_sensitivity": {k: float(f"{v:.3e}") for k, v in sens.items()},
           "bar": {k: float(f"{v:.3e}") for k, v in bar.items()},
           "within_bar": all(err[k] <= bar[k] for k in err), "oracle_s": round(cpu_s, 1)}
    print(json.dumps(out), flush=True)
    return 0 if out["within_bar"] and st_gpu == 0 and res[False][0] == 0 else 1


if __name__ == "__main__":
    sys.exit(main())
