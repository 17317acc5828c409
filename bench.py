"""Benchmark of the MI355X BioFVM backend.

Default workload (one GPU): the single large domain -- 200^3 voxels at dx=20um,
4 substrates with D log-uniform in [1,1e5] um^2/min and decay uniform in
[0.01,1]/min, dt=0.01 min, FP32 field uniform in [0,1], Dirichlet 1.0 on the
x-min face, bulk supply 0.5/min inside a central sphere of r=500um, 1e5 uniform
cells with secretion/uptake uniform in [0,10]/min -- stepped 1000 times.  A step
is one pass of the whole BioFVM step (three Thomas sweeps, Dirichlet, bulk,
cells) over the resident field; inputs are in HBM before the timed region.

Prints one JSON line: voxel*substrate updates/s, the roofline of the dominant
kernel (HIP events on the solver stream), and the CPU oracle timed on this
host's cores as a reported baseline.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "sweeps_200cube_s4_hbm_traffic.json")


def measured_traffic(kernel, side, S):
    """HBM bytes per launch of `kernel` (sweep_x/y/z) from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs of this workload, each counter
    scaled by what it reads for a copy of known size; None for other shapes."""
    if (side, S) != (200, 4) or not os.path.exists(TRAFFIC_PROFILE):
        return None
    prof = json.load(open(TRAFFIC_PROFILE))
    k = prof["kernels"].get(kernel)
    if k is None:
        return None
    kib = k["fetch_kib"] * prof["fetch_factor"] + k["write_kib"] * prof["write_factor"]
    return {"bytes": kib * 1024, "kernel_symbol": k["kernel"], "source": "profiles/" + prof["source"]}


def build_c2(side, S, n_cells, seed):
    """Config-2 inputs, generated without the oracle (same recipe as tests/scenarios.py)."""
    rng = np.random.default_rng(seed)
    dx = 20
    box = side * dx
    D = 10.0 ** rng.uniform(0.0, 5.0, S)
    lam = rng.uniform(0.01, 1.0, S)
    field = rng.uniform(0.0, 1.0, (side, side, side, S))
    c = (np.arange(side) + 0.5) * dx - box / 2.0
    zz, yy, xx = np.meshgrid(c, c, c, indexing="ij")
    inside = ((xx ** 2 + yy ** 2 + zz ** 2) <= 500.0 ** 2).reshape(-1)
    bulk_vox = np.nonzero(inside)[0]
    nb = len(bulk_vox)
    bulk = (bulk_vox, np.full((nb, S), 0.5), np.zeros((nb, S)), np.ones((nb, S)))
    pos = rng.uniform(0.0, box, (n_cells, 3))
    np.minimum(pos, np.nextafter(box, 0), out=pos)
    sec = rng.uniform(0.0, 10.0, (n_cells, S))
    upt = rng.uniform(0.0, 10.0, (n_cells, S))
    agents = dict(positions=pos, secretion=sec, saturation=np.ones((n_cells, S)), uptake=upt,
                  net_export=np.zeros((n_cells, S)), internalized=None,
                  fraction_released=np.zeros((n_cells, S)), volumes=np.full(n_cells, 2494.0),
                  is_active=np.ones(n_cells, dtype=np.uint8))
    return dict(side=side, S=S, D=D, lam=lam, field=field, bulk=bulk, agents=agents, dt=0.01, dx=dx)


def gpu_domain(w, precision="fp64"):
    from physicore_amd import bfvm

    side, S = w["side"], w["S"]
    d = bfvm.Domain(3, (side, side, side), (w["dx"],) * 3, w["dt"], w["D"], w["lam"], np.zeros(S),
                    precision=precision)
    d.upload_field(w["field"])
    vals = np.zeros((6, S))
    conds = np.zeros((6, S), dtype=np.uint8)
    vals[0], conds[0] = 1.0, 1
    d.set_dirichlet_faces(vals, conds)
    d.set_bulk(*w["bulk"])
    a = w["agents"]
    d.set_agents(a["positions"], a["secretion"], a["saturation"], a["uptake"], a["net_export"],
                 a["internalized"], a["fraction_released"], a["volumes"], a["is_active"])
    return d


def cpu_baseline(w, budget_s=12.0):
    """The CPU oracle (FP64 restatement of the OpenMP solver) on this host's cores,
    on a bounded sample of the same workload: as many whole steps as fit ~budget_s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc

    side, S = w["side"], w["S"]
    p = orc.Problem(3, (side, side, side), (w["dx"],) * 3, w["dt"], w["D"], w["lam"])
    field = w["field"].copy()
    dirs = orc.DirichletSpec(S, 3)
    dirs.set_face(0, "min", np.ones(S), np.ones(S, dtype=np.uint8))
    vox, sup, upt, tgt = w["bulk"]
    dense = [np.zeros(p.shape) for _ in range(3)]
    for arr, src in zip(dense, (sup, upt, tgt)):
        arr.reshape(-1, S)[vox] = src
    a = w["agents"]
    ag = orc.AgentSpec(len(a["volumes"]), S, 3)
    for k in ("positions", "secretion", "saturation", "uptake", "net_export", "fraction_released",
              "volumes", "is_active"):
        setattr(ag, k, np.ascontiguousarray(a[k]))
    s = orc.Solver(p, field, dirichlet=dirs, bulk=dense, agents=ag)
    s.solve(1)  # warm-up (includes the recompute), excluded like the reference benchmark
    steps, t0 = 0, time.perf_counter()
    while True:
        s.solve(1)
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s and steps >= 2:
            break
    ups = side ** 3 * S * steps / el
    return {"value": ups, "unit": "voxel*substrate updates/s", "cores": orc.num_threads(),
            "kind": "port",
            "sample": f"{steps} whole FP64 steps of the same {side}^3x{S} workload (incl. bulk, "
                      f"cells, Dirichlet) in {el:.1f}s, after 1 warm-up step"}


def build_c3_domains(seed=44, count=64, sides=tuple(range(32, 161, 16)), max_S=8, max_cells=50_000):
    """Batched ensemble: `count` domains, each axis side drawn from `sides`
    (anisotropic), dx=20, S uniform in 1..max_S with D/decay as the single large
    domain, field U[0,1], 0..max_cells uniform cells (secretion/uptake U[0,10],
    saturation 1, volume 2494).  Fields are generated and uploaded one domain at
    a time."""
    from physicore_amd import bfvm

    rng = np.random.default_rng(seed)
    doms, units = [], 0
    for _ in range(count):
        n = [int(rng.choice(sides)) for _ in range(3)]
        S = int(rng.integers(1, max_S + 1))
        D = 10.0 ** rng.uniform(0.0, 5.0, S)
        lam = rng.uniform(0.01, 1.0, S)
        d = bfvm.Domain(3, tuple(n), (20, 20, 20), 0.01, D, lam, np.zeros(S))
        d.upload_field(rng.uniform(0.0, 1.0, (n[2], n[1], n[0], S)).astype(np.float32))
        k = int(rng.integers(0, max_cells + 1))
        if k:
            box = np.array(n, dtype=np.float64) * 20.0
            pos = np.minimum(rng.uniform(0.0, 1.0, (k, 3)) * box, np.nextafter(box, 0))
            d.set_agents(pos, rng.uniform(0, 10, (k, S)), np.ones((k, S)), rng.uniform(0, 10, (k, S)),
                         np.zeros((k, S)), None, np.zeros((k, S)), np.full(k, 2494.0), np.ones(k, dtype=np.uint8))
        doms.append(d)
        units += n[0] * n[1] * n[2] * S
    return doms, units


def bench_batched(steps, graph_steps=10, seed=44):
    """The 64-domain ensemble stepped as one batch (after autotuning) and, for
    comparison, the same domains stepped one after another on the plain path."""
    from physicore_amd import bfvm

    doms, units = build_c3_domains(seed)
    b = bfvm.Batch(doms)
    tuning = b.autotune(steps=3)
    b.step(graph_steps, graph_steps)
    b.sync()
    t0 = time.perf_counter()
    b.step(steps, graph_steps)
    b.sync()
    el = time.perf_counter() - t0
    info = b.info()
    batched = units * steps / el
    # plain: each domain alone (default tuning), one after the other
    for d in doms:
        d.set_tuning()
        d.step_graph(graph_steps, graph_steps)
        d.sync()
    t0 = time.perf_counter()
    for d in doms:
        d.step_graph(steps, graph_steps)
        d.sync()
    el_plain = time.perf_counter() - t0
    plain = units * steps / el_plain
    gbs = units * 24 / (el / steps) / 1e9
    b.close()
    return {"workload": f"{len(doms)} domains, sides from {{32..160}} step 16 per axis, S 1-8, "
                        f"0-50k cells each, seed {seed}",
            "voxel_substrates": units, "steps": steps, "graph_steps": graph_steps,
            "value": batched, "unit": "voxel*substrate updates/s",
            "ms_per_step": el * 1e3 / steps,
            "plain_value": plain, "plain_ms_per_step": el_plain * 1e3 / steps,
            "batched_over_plain": batched / plain,
            "launches_per_step": info, "tuning": tuning,
            "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": gbs / HBM_PEAK_GBS, "bytes_per_step": units * 24}}


def bench_served(count, rate, seed=46):
    """Served stream: Poisson arrivals of requests (cube side log-uniform in
    [32, 256], S 1-6, 0-200k cells, 100-5000 steps), HBM budget 200 GB,
    continuous batching in chunks of 10 steps."""
    from physicore_amd import serve

    rng = np.random.default_rng(seed)
    reqs = serve.poisson_stream(rng, count, rate, random_field=False)
    srv = serve.Server(hbm_budget=200e9, chunk=10)
    for r in reqs:
        srv.submit(r)
    t0 = srv.now()
    done = srv.run()
    wall = srv.now() - t0
    ttc = [r.finished - r.arrival for r in done]
    return {"workload": f"{count} requests, Poisson rate {rate}/s, side log-uniform [32,256], S 1-6, "
                        f"0-200k cells, 100-5000 steps, budget 200 GB, chunk 10, seed {seed}",
            "requests": len(done), "wall_s": wall,
            "value": srv.updates / wall, "unit": "voxel*substrate updates/s",
            "p50_s": serve.percentile(ttc, 50), "p99_s": serve.percentile(ttc, 99),
            "chunks": srv.chunks}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--side", type=int, default=200)
    ap.add_argument("--substrates", type=int, default=4)
    ap.add_argument("--cells", type=int, default=100_000)
    ap.add_argument("--graph-steps", type=int, default=10)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--batched-steps", type=int, default=100)
    ap.add_argument("--served-requests", type=int, default=32)
    ap.add_argument("--served-rate", type=float, default=8.0)
    ap.add_argument("--only-plain", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy a 256 MB device buffer 10 times (reference for HBM counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch

        src = torch.rand(64 * 1024 * 1024, device="cuda:0")  # 256 MB of FP32
        dst = torch.empty_like(src)
        for _ in range(10):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": src.numel() * 4, "copy_bytes_written": dst.numel() * 4,
                          "launches": 10}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; --gpus above 1 is not supported "
              "(independent domains would spread over GPUs by HBM budget, one large domain by z-slabs)",
              file=sys.stderr)
        sys.exit(2)

    from physicore_amd import bfvm

    if bfvm.device_count() < 1:
        print("bench.py: no HIP device", file=sys.stderr)
        sys.exit(1)
    w = build_c2(args.side, args.substrates, args.cells, seed=43)
    d = gpu_domain(w, args.precision)
    V = args.side ** 3
    units = V * args.substrates

    d.step_graph(args.warmup, args.graph_steps)
    d.sync()
    t0 = time.perf_counter()
    d.step_graph(args.steps, args.graph_steps)
    d.sync()
    el = time.perf_counter() - t0
    ms_per_step = el * 1e3 / args.steps
    value = units * args.steps / el

    # per-stage launch durations with HIP events on the solver stream (each
    # launch timed on its own, so each figure includes the gap before it),
    # then each sweep's average over 50 back-to-back launches between two
    # events, as the kernels run inside the step graph (the extra sweeps only
    # change the field after the timed region)
    stages = d.time_stages(min(args.steps, 50))
    sweeps = {a: d.time_sweep(i, 50) for i, a in enumerate("xyz")}
    sweep_bytes = units * 8  # one sweep reads and writes every FP32 value once
    dominant = max(sweeps, key=sweeps.get)
    achieved = sweep_bytes / (sweeps[dominant] * 1e-6) / 1e9
    diff_us = sweeps["x"] + sweeps["y"] + sweeps["z"]
    traffic = measured_traffic(f"sweep_{dominant}", args.side, args.substrates)
    step_gbs = units * 24 / (ms_per_step * 1e-3) / 1e9

    out = {
        "metric": "voxel*substrate updates/s",
        "value": value,
        "unit": "voxel*substrate updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 field, f64 sweep arithmetic" if args.precision == "fp64" else "f32",
        "data": "synthetic",
        "config": {"workload": f"single large domain {args.side}^3 voxels x {args.substrates} substrates, "
                               f"dx=20um, dt=0.01min, x-min Dirichlet, bulk sphere r=500um, "
                               f"{args.cells} cells, whole BioFVM step",
                   "graph_steps": args.graph_steps, "seed": 43},
        "roofline": {"bound": "hbm", "kernel": f"sweep_{dominant}", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic["bytes"] if traffic else None,
                     "traffic_source": traffic,
                     "algorithmic_bytes_per_launch": sweep_bytes,
                     "avg_launch_us": sweeps[dominant],
                     "avg_launch_us_timed_alone": stages[dominant]},
        "step_roofline": {"bytes_per_step": units * 24, "achieved": step_gbs,
                          "frac": step_gbs / HBM_PEAK_GBS,
                          "diffusion_us": diff_us,
                          "diffusion_frac": units * 24 / (diff_us * 1e-6) / 1e9 / HBM_PEAK_GBS},
        "stages_us": stages,
        "sweeps_back_to_back_us": sweeps,
    }
    del d
    if not args.only_plain:
        print("bench.py: batched ensemble", file=sys.stderr, flush=True)
        out["batched"] = bench_batched(args.batched_steps)
        print("bench.py: served stream", file=sys.stderr, flush=True)
        out["served"] = bench_served(args.served_requests, args.served_rate)
    if not args.no_cpu_baseline:
        print("bench.py: CPU baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(w, args.cpu_budget)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
