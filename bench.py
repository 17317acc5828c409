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

from physicore_amd import workloads as wl  # noqa: E402

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


def _oracle_inputs(w):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import scenarios

    return scenarios.oracle_inputs(w)


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc

    return orc


def cpu_baseline(w, budget_s=12.0, recompute_every=0):
    """The CPU oracle (FP64 restatement of the OpenMP solver) on this host's cores,
    on a bounded sample of the same workload: as many whole steps as fit ~budget_s
    (with a positional recompute every `recompute_every` steps when asked)."""
    orc = _oracle()
    p, field, dirs, bulk, ag = _oracle_inputs(w)
    s = orc.Solver(p, field, dirichlet=dirs, bulk=bulk, agents=ag, compute_internalized=w.compute_internalized)
    s.solve(1)  # warm-up (includes the recompute), excluded like the reference benchmark
    steps, t0 = 0, time.perf_counter()
    while True:
        if recompute_every and steps % recompute_every == 0:
            s.recompute()
        s.solve(1)
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s and steps >= 2:
            break
    ups = w.units * steps / el
    rc = f", recompute every {recompute_every} steps" if recompute_every and "recompute" not in w.name else ""
    return {"value": ups, "unit": "voxel*substrate updates/s", "cores": orc.num_threads(),
            "kind": "port",
            "sample": f"{steps} whole FP64 steps of the same workload ({w.name}{rc}) in {el:.1f}s, "
                      f"after 1 warm-up step"}


def cpu_baseline_domains(workloads, steps_each=2):
    """The CPU oracle over several domains one after another (the sequential CPU
    form of a batch or a request stream): per domain one warm-up step, then
    `steps_each` timed steps.  Returns the aggregate updates/s and the per-domain
    seconds per step."""
    orc = _oracle()
    units = secs = 0.0
    per_step = []
    for w in workloads:
        p, field, dirs, bulk, ag = _oracle_inputs(w)
        s = orc.Solver(p, field, dirichlet=dirs, bulk=bulk, agents=ag,
                       compute_internalized=w.compute_internalized)
        s.solve(1)
        t0 = time.perf_counter()
        s.solve(steps_each)
        el = time.perf_counter() - t0
        units += w.units * steps_each
        secs += el
        per_step.append(el / steps_each)
    return units / secs, per_step, orc.num_threads()


def build_c3_domains(seed=44, count=64):
    """The batched ensemble (physicore_amd.workloads.c3): `count` domains, each
    axis side drawn from {32..160} (anisotropic), dx=20, S uniform in 1..8 with
    D/decay as the single large domain, field U[0,1], 0..50k uniform cells.
    Fields are generated and uploaded one domain at a time."""
    doms, units = [], 0
    for w in wl.c3(seed, count):
        doms.append(wl.make_domain(w))
        units += w.units
    return doms, units


def bench_batched(steps, graph_steps=10, seed=44, cpu=True, compare=True):
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
    if not compare:
        b.close()
        return {"workload": f"{len(doms)} domains, seed {seed}", "voxel_substrates": units, "steps": steps,
                "value": batched, "ms_per_step": el * 1e3 / steps, "launches_per_step": info, "tuning": tuning}
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
    count = len(doms)
    del doms
    out = {"workload": f"{count} domains, sides from {{32..160}} step 16 per axis, S 1-8, "
                        f"0-50k cells each, seed {seed}",
            "voxel_substrates": units, "steps": steps, "graph_steps": graph_steps,
            "value": batched, "unit": "voxel*substrate updates/s",
            "ms_per_step": el * 1e3 / steps,
            "plain_value": plain, "plain_ms_per_step": el_plain * 1e3 / steps,
            "batched_over_plain": batched / plain,
            "launches_per_step": info, "tuning": tuning,
            "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": gbs / HBM_PEAK_GBS, "bytes_per_step": units * 24}}
    if cpu:
        print("bench.py: CPU baseline of the ensemble", file=sys.stderr, flush=True)
        ups, _, cores = cpu_baseline_domains(wl.c3(seed), steps_each=2)
        out["cpu_baseline"] = {"value": ups, "unit": "voxel*substrate updates/s", "cores": cores, "kind": "port",
                               "sample": "the FP64 oracle over all 64 domains one after another, "
                                         "2 timed steps each after 1 warm-up step"}
    return out


def bench_served(count, rate, seed=46, cpu=True, **server_kw):
    """Served stream: Poisson arrivals of requests (cube side log-uniform in
    [32, 256], S 1-6, 0-200k cells, 100-5000 steps), HBM budget 200 GB,
    continuous batching in chunks of 10 steps."""
    from physicore_amd import serve

    rng = np.random.default_rng(seed)
    reqs = serve.poisson_stream(rng, count, rate, random_field=False)
    srv = serve.Server(**{"hbm_budget": 200e9, "chunk": 10, **server_kw})
    for r in reqs:
        srv.submit(r)
    t0 = srv.now()
    done = srv.run()
    wall = srv.now() - t0
    srv.close()
    ttc = [r.finished - r.arrival for r in done]
    out = {"workload": f"{count} requests, Poisson rate {rate}/s, side log-uniform [32,256], S 1-6, "
                        f"0-200k cells, 100-5000 steps, budget 200 GB, chunk 10, seed {seed}",
            "requests": len(done), "wall_s": wall,
            "value": srv.updates / wall, "unit": "voxel*substrate updates/s",
            "p50_s": serve.percentile(ttc, 50), "p99_s": serve.percentile(ttc, 99),
            "chunks": srv.chunks}
    # each request alone on the idle GPU (graphs of 10 steps): the floor of its
    # time-to-completion, and their sum, the stream's work done sequentially
    alone = {}
    for r in done:
        d = serve.domain_for(r)
        d.sync()
        t1 = time.perf_counter()
        g = r.steps - r.steps % 10
        if g:
            d.step_graph(g, 10)
        d.step(r.steps - g)
        d.sync()
        alone[r.rid] = time.perf_counter() - t1
        d.close()
    slow = [(r.finished - r.arrival) / alone[r.rid] for r in done]
    out["alone"] = {"sequential_s": float(sum(alone.values())),
                    "p50_s": serve.percentile(list(alone.values()), 50),
                    "p99_s": serve.percentile(list(alone.values()), 99),
                    "max_s": max(alone.values()),
                    "ttc_over_alone_p50": serve.percentile(slow, 50),
                    "ttc_over_alone_p99": serve.percentile(slow, 99),
                    "ttc_of_longest_alone": next(r.finished - r.arrival for r in done
                                                 if alone[r.rid] == max(alone.values()))}
    if cpu:
        # the oracle per request, one after another: 2 timed steps each, scaled
        # to the request's step count
        print("bench.py: CPU baseline of the stream", file=sys.stderr, flush=True)
        reqs = serve.poisson_stream(np.random.default_rng(seed), count, rate, random_field=False)
        ups, per_step, cores = cpu_baseline_domains([serve.workload_of(r) for r in reqs], steps_each=2)
        t_req = [t * r.steps for t, r in zip(per_step, reqs)]
        out["cpu_baseline"] = {"value": ups, "unit": "voxel*substrate updates/s", "cores": cores, "kind": "port",
                               "sequential_s": float(sum(t_req)),
                               "p50_request_s": serve.percentile(t_req, 50),
                               "p99_request_s": serve.percentile(t_req, 99),
                               "sample": f"the FP64 oracle over all {count} requests one after another, 2 timed "
                                         f"steps each after 1 warm-up step, scaled to each request's steps"}
    return out


def bench_dense(steps=200, cpu=True, budget_s=12.0):
    """The dense-agent workload (workloads.c4: 400x400x100 x 2 substrates,
    2e6 cells in a Gaussian spheroid, net export), stepped in 10-step graphs
    with the atomic and with the sorted (bit-reproducible) cell reduction."""
    w = wl.c4()
    out = {"workload": w.name, "steps": steps, "voxel_substrates": w.units, "unit": "voxel*substrate updates/s"}
    for reduce in ("atomic", "sorted"):
        d = wl.make_domain(w, reduce=reduce)
        d.step_graph(20, 10)
        d.sync()
        t0 = time.perf_counter()
        d.step_graph(steps, 10)
        d.sync()
        el = time.perf_counter() - t0
        out[reduce] = {"value": w.units * steps / el, "ms_per_step": el * 1e3 / steps}
        if reduce == "atomic":
            out["schedule"] = d.sweep_plan()
            out["stages_us"] = d.time_stages(30)
        d.close()
    out["value"] = out["atomic"]["value"]
    if cpu:
        print("bench.py: CPU baseline of the dense-agent workload", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(w, budget_s)
    return out


def bench_reference(steps=100, cpu=True, budget_s=12.0):
    """The reference's own benchmark workload (openmp_solver/examples/benchmark.cpp:
    250^3 x 4, bulk in every voxel, six Dirichlet faces, 2e6+1 lattice agents,
    internalized tracking), stepped with a positional recompute every 10 steps
    as that program does, and without."""
    w = wl.reference_benchmark(steps=steps)
    d = wl.make_domain(w)
    period = w.recompute_every
    d.step_graph(period, period)
    d.sync()
    t0 = time.perf_counter()
    for _ in range(steps // period):
        d.recompute_positional_data()
        d.step_graph(period, period)
    d.sync()
    el_rc = time.perf_counter() - t0
    t0 = time.perf_counter()
    d.step_graph(steps, period)
    d.sync()
    el = time.perf_counter() - t0
    stages = d.time_stages(period)
    out = {"workload": w.name, "steps": steps, "voxel_substrates": w.units,
           "value": w.units * steps / el_rc, "unit": "voxel*substrate updates/s",
           "ms_per_step_with_recompute": el_rc * 1e3 / steps,
           "ms_per_step_without_recompute": el * 1e3 / steps,
           "stages_us": stages}
    d.close()
    if cpu:
        print("bench.py: CPU baseline of the reference workload", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(w, budget_s, recompute_every=period)
    return out


def bench_plain(args):
    """The single large domain (workloads.c2), whole steps replayed from
    graphs; returns the workload and the headline fields of the JSON line."""
    w = wl.c2(args.side, args.substrates, args.cells, seed=43)
    d = wl.make_domain(w, precision=args.precision)
    units = w.units

    d.step_graph(args.warmup, args.graph_steps)
    d.sync()
    t0 = time.perf_counter()
    d.step_graph(args.steps, args.graph_steps)
    d.sync()
    el = time.perf_counter() - t0
    ms_per_step = el * 1e3 / args.steps
    value = units * args.steps / el

    # in-step launch durations: 50 steps queued back to back on the solver
    # stream with HIP events between the launches (each figure includes the
    # gap before its launch; these steps launch each sweep once, while the
    # timed steps above run the x sweep in two halves on two streams, the
    # second overlapping the y lines of the first half's planes); the
    # roofline is priced on these.  Then, as a
    # kernel-level figure, each sweep's average over 50 launches of that sweep
    # alone between two events (the extra sweeps only change the field after
    # the timed region)
    plan = d.sweep_plan()
    schedule = dict(plan, note=(
        "x sweep in two halves on two streams (split_x and x_tiles - split_x tiles); the y tiles of the planes "
        "the first half covers (split_y) start when it is done, the others after both halves; z after all of "
        "y; results bit-identical to one launch per sweep" if plan["split_x"] else "one launch per sweep"))
    stages = d.time_stages(min(args.steps, 50))
    sweeps = {a: d.time_sweep(i, 50) for i, a in enumerate("xyz")}
    sweep_bytes = units * 8  # one sweep reads and writes every FP32 value once
    dominant = max("xyz", key=lambda a: stages[a])
    achieved = sweep_bytes / (stages[dominant] * 1e-6) / 1e9
    diff_us = stages["x"] + stages["y"] + stages["z"]
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
                   "graph_steps": args.graph_steps, "seed": 43,
                   "precision": ("the field is stored in FP32 (stochastic rounding of every store), the "
                                 "Thomas recurrences run in FP64; the reference keeps the field in FP64 "
                                 "(real_t = double); results are held to 1e-5 relative L-inf of the FP64 "
                                 "oracle over 1000 steps" if args.precision == "fp64" else
                                 "FP32 field and FP32 recurrences (reduced precision)")},
        "roofline": {"bound": "hbm", "kernel": f"sweep_{dominant}", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic["bytes"] if traffic else None,
                     "traffic_source": traffic,
                     "algorithmic_bytes_per_launch": sweep_bytes,
                     "avg_launch_us": stages[dominant],
                     "timing": "in-step: HIP events between the launches of 50 back-to-back steps "
                               "with one launch per sweep",
                     "kernel_level": {"avg_launch_us": sweeps[dominant],
                                      "achieved": sweep_bytes / (sweeps[dominant] * 1e-6) / 1e9,
                                      "timing": "50 launches of the sweep alone between two events"}},
        "schedule": schedule,
        "step_roofline": {"bytes_per_step": units * 24, "achieved": step_gbs,
                          "frac": step_gbs / HBM_PEAK_GBS,
                          "diffusion_us": diff_us,
                          "diffusion_frac": units * 24 / (diff_us * 1e-6) / 1e9 / HBM_PEAK_GBS},
        "stages_us": stages,
        "sweeps_back_to_back_us": sweeps,
    }
    d.close()
    return w, out


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
    ap.add_argument("--only-plain", action="store_true", help="same as --legs plain")
    ap.add_argument("--legs", default="plain,batched,served,reference",
                    help="comma-separated parts to run: plain, batched, served, reference, dense")
    ap.add_argument("--dense-steps", type=int, default=200)
    ap.add_argument("--reference-steps", type=int, default=100)
    ap.add_argument("--batched-no-compare", action="store_true",
                    help="batched leg without the one-domain-at-a-time comparison (for kernel traces)")
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
    legs = {"plain"} if args.only_plain else set(args.legs.split(","))
    out = {}
    if "plain" in legs:
        w, out = bench_plain(args)
    cpu = not args.no_cpu_baseline
    if "batched" in legs:
        print("bench.py: batched ensemble", file=sys.stderr, flush=True)
        out["batched"] = bench_batched(args.batched_steps, cpu=cpu, compare=not args.batched_no_compare)
    if "served" in legs:
        print("bench.py: served stream", file=sys.stderr, flush=True)
        out["served"] = bench_served(args.served_requests, args.served_rate, cpu=cpu)
    if "reference" in legs:
        print("bench.py: reference benchmark workload", file=sys.stderr, flush=True)
        out["reference_benchmark"] = bench_reference(args.reference_steps, cpu=cpu, budget_s=args.cpu_budget)
    if "dense" in legs:
        print("bench.py: dense-agent workload", file=sys.stderr, flush=True)
        out["dense_agents"] = bench_dense(args.dense_steps, cpu=cpu, budget_s=args.cpu_budget)
    if cpu and "plain" in legs:
        print("bench.py: CPU baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(w, args.cpu_budget)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
