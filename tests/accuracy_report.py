"""Error of the GPU path against the FP64 oracle for every workload, in both
recurrence precisions (FP32 field in both), plus the per-substrate mass
balance.  Not a pytest module: run it on a machine with a GPU,
`python tests/accuracy_report.py [--quick]`; it prints one JSON line per case.
Long cases run a bounded number of steps (stated in each line)."""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(HERE))

import oracle_path  # noqa: F401,E402  (puts the oracle on sys.path)
import oracle as orc  # noqa: E402
import scenarios as sc  # noqa: E402
from gpu_helpers import gpu_domain, rel_linf  # noqa: E402


def run_case(name, scn, steps, internalized=False, reduces=("atomic", "sorted"), tunings=(None,)):
    p, field, dirs, bulk, ag = scn[:5]
    f0 = field.astype(np.float32).astype(np.float64)
    t0 = time.perf_counter()
    want = f0.copy()
    s = orc.Solver(p, want, dirichlet=dirs, bulk=list(bulk) if bulk is not None else None, agents=ag,
                   compute_internalized=internalized)
    s.solve(steps)
    t_cpu = time.perf_counter() - t0
    out = {"case": name, "steps": steps, "grid": [p.nx, p.ny, p.nz], "S": p.S,
           "cells": 0 if ag is None else int(ag.n), "oracle_s": round(t_cpu, 2)}
    V = float(np.prod(p.voxel))
    for k, tuning in enumerate(tunings):
        tag = "" if tuning is None else f"_t{k}"
        if tuning is not None:
            out[f"tuning{tag}"] = tuning
        for prec in ("fp64", "fp32"):
            for red in reduces:
                d = gpu_domain(p, field=field, dirichlet=dirs, bulk=bulk, agents=ag, internalized=internalized,
                               precision=prec, reduce=red)
                if tuning:
                    d.set_tuning(**tuning)
                d.step_graph(steps, min(10, steps))
                got = d.field()
                out[f"{prec}_{red}_rel_linf{tag}"] = [float(x) for x in rel_linf(got, want)]
                if prec == "fp64" and red == reduces[0]:
                    # where the largest error sits: (z, y, x) and the FP64 value there
                    err = np.abs(got.reshape(-1, p.S) - want.reshape(-1, p.S))
                    at = err.argmax(axis=0)
                    out[f"worst_voxel_zyx{tag}"] = [[int(i // (p.nx * p.ny)), int(i // p.nx % p.ny),
                                                     int(i % p.nx)] for i in at]
                    out[f"worst_value{tag}"] = [float(want.reshape(-1, p.S)[i, s]) for s, i in enumerate(at)]
                    out[f"mass_gpu{tag}"] = [float(x) for x in d.total_mass()]
    out["mass_oracle"] = [float(x) for x in (want.reshape(-1, p.S).sum(axis=0) * V)]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--quick", action="store_true", help="fewer steps on the large cases")
    ap.add_argument("--only", default="", help="run only the cases whose name contains this text")
    ap.add_argument("--tuning", action="append", default=[],
                    help='kernel selection for the GPU runs, as JSON (keyword arguments of '
                    'Domain.set_tuning), e.g. \'{"round_nearest": 1}\'; may be repeated, "null" = the default')
    a = ap.parse_args()
    q = a.quick
    cases = []
    cases.append(("config1 50^3x1, 1000 cells", sc.config1(), 100))
    p, field, dirs, bulk, ag = sc.config2(side=200, S=4, n_cells=100_000)
    cases.append(("config2 200^3x4, 1e5 cells", (p, field, dirs, bulk, ag), 20 if q else 1000))
    rng = np.random.default_rng(44)
    for k in range(3):
        scn = sc.mixed_domain(rng, list(range(32, 161, 16)), max_S=8, max_cells=50_000)
        cases.append((f"config3 domain {k} {scn[0].nx}x{scn[0].ny}x{scn[0].nz}x{scn[0].S}", scn[:5],
                      20 if q else 500))
    cases.append(("config4 400x400x100x2, 2e6 cells", sc.config4(), 3 if q else 200))
    # the served stream of the benchmark (seed 46): its requests of side <= 96
    # (the oracle finishes them in seconds) at their own step counts
    from physicore_amd import serve
    from physicore_amd import workloads as wl

    reqs = serve.poisson_stream(np.random.default_rng(46), 32, 8.0, random_field=False)
    for r in [r for r in reqs if r.grid[0] <= 96][:3]:
        w = serve.workload_of(r)
        cases.append((f"config5 request {r.rid} {r.grid[0]}^3x{r.S}, {w.n_agents} cells", sc.oracle_inputs(w),
                      min(r.steps, 50) if q else r.steps))
    refw = wl.reference_benchmark()
    cases.append(("reference benchmark 250^3x4, 2e6+1 agents, internalized", sc.oracle_inputs(refw),
                  10 if q else 100, True))
    for name, scn, steps, *internalized in cases:
        if a.only in name:
            tunings = [json.loads(t) for t in a.tuning] or [None]
            print(json.dumps(run_case(name, scn, steps, internalized=bool(internalized and internalized[0]),
                                      tunings=tunings)), flush=True)


if __name__ == "__main__":
    main()
