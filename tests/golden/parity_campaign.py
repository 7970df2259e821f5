"""Parity campaign on the UQ ensemble itself: the GPU batched FP64 path
against the CPU oracle on every member of the 256-member ensemble (seed
2026) with a given side, over a given number of steps (default: the 64 x 64
members over their full 5000 steps).

Not a pytest module (it takes minutes of CPU). Run from the repository root
on a machine with the GPU and the built library:

    python tests/golden/parity_campaign.py --side 64 --steps 5000 --workers 16

The oracle runs first, one member per worker process (nothing touches the
GPU before the pool has finished); then one Batch steps all selected
members together, as the headline benchmark does. For every member (or the
first `--floor-members`) the oracle is also run from a one-ulp perturbed
initial state (the same perturbation in every column: the members are
uniform along x), which gives the oracle's own sensitivity over the run.

Bar per field, as for the golden fixtures: relative L2 <= max(1e-12, 8 x
that sensitivity), or 1e-12 for a member without the sensitivity run. The shear-wave members need the floor: their velocity
decays viscously by up to ~1e-6 over 5000 steps, so its relative L2 is
ill-conditioned (the oracle's own one-ulp sensitivity reaches ~1e-8); the
error against the initial amplitude is reported beside it.

Prints one JSON line per member and a summary line; with --out also writes
them as JSON.
"""
from __future__ import annotations

import argparse
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, HERE)

from uq_amd import cases  # noqa: E402
import make_golden  # noqa: E402


def _oracle_job(job):
    index, steps, perturb = job
    case = cases.uq_member(index, steps=steps)
    t = time.time()
    st, m = make_golden.run_oracle(case, steps, perturb=perturb, x_invariant=True)
    return index, perturb, st, {k: m[k] for k in make_golden.FIELDS}, time.time() - t


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--side", type=int, default=64)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--workers", type=int, default=16)
    ap.add_argument("--floor-members", type=int, default=-1, help="-1: all members")
    ap.add_argument("--limit", type=int, default=0, help="at most this many members (0: all)")
    ap.add_argument("--out", default="")
    a = ap.parse_args()

    ens = cases.uq_ensemble(256, seed=2026, steps=a.steps)
    idx = [i for i, c in enumerate(ens) if c.nx == a.side]
    if a.limit:
        idx = idx[:a.limit]
    kinds = {i: ens[i].meta["kind"] for i in idx}
    floor_idx = idx if a.floor_members < 0 else idx[:a.floor_members]
    jobs = [(i, a.steps, False) for i in idx] + [(i, a.steps, True) for i in floor_idx]
    print(f"[campaign] {len(idx)} members of side {a.side}, {a.steps} steps; "
          f"{len(floor_idx)} sensitivity runs; {a.workers} workers", flush=True)
    t0 = time.time()
    ref, pert = {}, {}
    with mp.get_context("fork").Pool(a.workers) as pool:
        for n, (i, p, st, m, sec) in enumerate(pool.imap_unordered(_oracle_job, jobs), 1):
            (pert if p else ref)[i] = (st, m)
            if n % 8 == 0 or n == len(jobs):
                print(f"[campaign] oracle {n}/{len(jobs)} done ({time.time() - t0:.0f} s)", flush=True)
    cpu_s = time.time() - t0

    from uq_amd import lbm  # the GPU is first touched here, after the pool
    solvers = [lbm.Solver(cases.uq_member(i, steps=a.steps), "fp64") for i in idx]
    initial = [s.get_macro() for s in solvers]
    b = lbm.Batch(solvers)
    t1 = time.time()
    st_gpu = b.step(a.steps)
    gpu_s = time.time() - t1
    rows = []
    u = lambda d: np.stack([d["ux"], d["uy"]])
    for i, s, m0 in zip(idx, solvers, initial):
        m = s.get_macro()
        info = s.info()
        st_o, mo = ref[i]
        err = make_golden.field_errors(m, mo)
        row = {"member": i, "kind": kinds[i], "oracle_status": st_o, "gpu_step": info["step"],
               "errors": {k: float(f"{v:.3e}") for k, v in err.items()},
               "u_decay": float(f"{np.linalg.norm(u(mo)) / np.linalg.norm(u(m0)):.3e}"),
               "u_error_vs_initial": float(f"{np.linalg.norm(u(m) - u(mo)) / np.linalg.norm(u(m0)):.3e}")}
        if i in pert:
            sens = make_golden.field_errors(pert[i][1], mo)
            row["oracle_1ulp_sensitivity"] = {k: float(f"{v:.3e}") for k, v in sens.items()}
            row["within_bar"] = all(err[k] <= max(1e-12, 8.0 * sens[k]) for k in err)
        else:  # no sensitivity run: the plain 1e-12 bar
            row["within_bar"] = all(v <= 1e-12 for v in err.values())
        rows.append(row)
        print(json.dumps(row), flush=True)
    b.close()
    for s in solvers:
        s.close()
    worst = {k: max(r["errors"][k] for r in rows) for k in ("rho", "u", "T")}
    floor = {k: max(r["oracle_1ulp_sensitivity"][k] for r in rows if "oracle_1ulp_sensitivity" in r)
             for k in ("rho", "u", "T")} if pert else {}
    summary = {"side": a.side, "steps": a.steps, "members": len(idx), "batch_status": st_gpu,
               "worst_rel_l2": worst, "worst_oracle_1ulp_sensitivity": floor,
               "worst_u_error_vs_initial": max(r["u_error_vs_initial"] for r in rows),
               "members_checked_against_bar": sum("within_bar" in r for r in rows),
               "members_within_bar": sum(bool(r.get("within_bar")) for r in rows),
               "oracle_cpu_s": round(cpu_s, 1), "oracle_workers": a.workers, "gpu_s": round(gpu_s, 2),
               "all_oracle_ok": all(r["oracle_status"] == 0 for r in rows)}
    print(json.dumps({"summary": summary}), flush=True)
    if a.out:
        with open(a.out, "w") as fh:
            json.dump({"summary": summary, "members": rows}, fh, indent=1)
    ok = summary["all_oracle_ok"] and st_gpu == 0 and \
        summary["members_within_bar"] == summary["members_checked_against_bar"]
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
