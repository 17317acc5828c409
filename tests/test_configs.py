"""Every benchmark workload at its own size, on the GPU, against the FP64 oracle.

The bar is the benchmark's: at most 1e-5 relative L-infinity error per
substrate after the workload's steps (FP32 field against the FP64 oracle run of
the same method); a single step or a few steps are held to 2e-6.  Total mass
per substrate is held to 1e-7 relative of the oracle's.  Cell sums are FP64 on
both sides: the sorted reduction is reproducible bit for bit run to run, the
atomic one matches it to 1e-13.  Batched and served results equal the same
domain stepped alone bit for bit (sorted reduction).

Workloads (physicore_amd/workloads.py):
  c2                   200^3 x 4, 1e5 cells, the full 1000 steps
  c3                   the 64-domain mixed ensemble (sides {32..160}, S 1-8, 0-50k cells)
  c4                   400x400x100 x 2 at dx=10, 2e6 clustered cells
  served stream        Poisson requests, sides 32-256, S 1-6, 0-200k cells
  reference benchmark  250^3 x 4, bulk in every voxel, six Dirichlet faces, 2e6+1
                       lattice agents, internalized tracking, recompute every 10 steps
"""
import sys

import numpy as np
import pytest

import oracle as orc
import scenarios as sc
from gpu_helpers import rel_linf, require_gpu
from physicore_amd import bfvm, serve
from physicore_amd import workloads as wl

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _gpu():
    require_gpu()


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


def _oracle_run(w, steps, quirks=False, recompute_every=0):
    """The oracle's field after `steps` steps of workload w, started from the
    FP32-rounded field the GPU holds; also the Solver (for its agents)."""
    p, field, dirs, bulk, ag = sc.oracle_inputs(w, quirks=quirks)
    f = field.astype(np.float32).astype(np.float64)
    s = orc.Solver(p, f, dirichlet=dirs, bulk=bulk, agents=ag, compute_internalized=w.compute_internalized)
    done = 0
    while done < steps:
        k = steps - done if not recompute_every else min(recompute_every, steps - done)
        if recompute_every and done:
            s.recompute()
        s.solve(k)
        done += k
    return f, s, p


def _voxel_of(w, positions):
    """Linear voxel index of each position (truncating, as cartesian_mesh::voxel_position)."""
    idx = (np.asarray(positions) / np.asarray(w.voxel, dtype=np.float64)).astype(np.int64)
    return (idx[:, 2] * w.grid[1] + idx[:, 1]) * w.grid[0] + idx[:, 0]


def _mass(field, w):
    return field.reshape(-1, w.S).sum(axis=0) * float(np.prod(w.voxel))


# ------------------------------------------------------------------ c2

def test_c2_full_1000_steps():
    """The single large domain for its full 1000 steps (10-step graphs, as the
    benchmark runs it): <= 1e-5 per substrate, mass within 1e-7."""
    w = wl.c2()
    _say(f"c2: {w.name}, {w.steps} steps")
    d = wl.make_domain(w)
    d.step_graph(w.steps, 10)
    want, _, _ = _oracle_run(w, w.steps)
    got = d.field()
    err = rel_linf(got, want)
    _say(f"c2: rel L-inf per substrate {err}")
    assert err.max() < 1e-5, err
    m_gpu, m_or = d.total_mass(), _mass(want, w)
    assert np.all(np.abs(m_gpu - m_or) <= 1e-7 * np.abs(m_or)), (m_gpu, m_or)


# ------------------------------------------------------------------ c4

@pytest.fixture(scope="module")
def c4_workload():
    return wl.c4()


def test_c4_full_size_atomic_and_sorted(c4_workload):
    """400x400x100x2 with 2e6 Gaussian cells through the default kernels
    (two-lane chunked x rows of 400, twisted y lines of 400), 10 steps, both
    reductions against the oracle; the sorted path is reproducible bit for bit
    and the atomic sums match the sorted ones to 1e-13."""
    w = c4_workload
    steps = 10
    want, s, _ = _oracle_run(w, steps)
    fields, sums = {}, {}
    for reduce in ("atomic", "sorted", "sorted"):
        d = wl.make_domain(w, reduce=reduce)
        d.step_graph(steps, 10)
        got = d.field()
        err = rel_linf(got, want)
        _say(f"c4 {reduce}: rel L-inf per substrate {err}")
        assert err.max() < 2e-6, (reduce, err)
        if reduce in fields:
            assert np.array_equal(fields[reduce], got), "sorted reduction not reproducible"
        fields[reduce] = got
        sums[reduce] = d.cell_sums()
        d.close()
    # owners are elected differently (first to win the ballot, or the first
    # agent of the voxel): compare the sums voxel by voxel
    by_voxel = {}
    for reduce in ("atomic", "sorted"):
        owners, num, den, fac, conflict = sums[reduce]
        vox = _voxel_of(w, w.agents["positions"][owners])
        order = np.argsort(vox)
        by_voxel[reduce] = (vox[order], num[order], den[order], fac[order], conflict)
    va, na, da, fa, ca = by_voxel["atomic"]
    vs, ns, ds, fs, cs = by_voxel["sorted"]
    assert np.array_equal(va, vs) and ca == cs
    for a, b in ((na, ns), (da, ds), (fa, fs)):
        assert np.all(np.abs(a - b) <= 1e-13 * np.maximum(np.abs(b), 1.0))


def test_c4_internalized_tracking(c4_workload):
    """The same workload with internalized tracking (many shared voxels, so the
    per-agent update runs): field and internalized amounts against the oracle."""
    w = c4_workload
    w2 = wl.Workload(**{**w.__dict__, "compute_internalized": True})
    steps = 5
    want, s, _ = _oracle_run(w2, steps)
    d = wl.make_domain(w2, reduce="sorted")
    d.step(steps)
    assert rel_linf(d.field(), want).max() < 2e-6
    err_i = rel_linf(d.internalized(), s.agents.internalized)
    _say(f"c4 internalized: rel L-inf {err_i}")
    assert err_i.max() < 2e-6


# ------------------------------------------------------------------ c3

def _c3_pair(k):
    """Domain k of the ensemble twice (one for the batch, one alone), from one
    generated field."""
    w = wl.c3_domain(k)
    return w, wl.make_domain(w, reduce="sorted"), wl.make_domain(w, reduce="sorted")


def test_c3_ensemble_batch_equals_alone_and_oracle():
    """The 64-domain ensemble of the benchmark: 20 steps as one batch (two
    10-step graphs) give every domain exactly its result alone; four domains
    (S = 5, 7, 8 and the smallest S) against the oracle."""
    steps = 20
    alone, mixed, specs = [], [], []
    for k in range(64):
        w, a, m = _c3_pair(k)
        specs.append(w)
        alone.append(a)
        mixed.append(m)
    _say(f"c3: {len(specs)} domains, {sum(w.units for w in specs)} voxel*substrates")
    b = bfvm.Batch(mixed)
    b.step(steps, 10)
    b.sync()
    for d in alone:
        d.step_graph(steps, 10)
    for k, (a, m) in enumerate(zip(alone, mixed)):
        assert np.array_equal(a.field(), m.field()), specs[k].name
    b.close()
    picks = []
    for pred in (lambda S: S == 5, lambda S: S == 7, lambda S: S == 8, lambda S: S <= 2):
        k = next(i for i, w in enumerate(specs) if pred(w.S) and i not in picks)
        picks.append(k)
    for k in picks:
        w = wl.c3_domain(k)
        want, _, _ = _oracle_run(w, steps)
        err = rel_linf(mixed[k].field(), want)
        _say(f"c3 {w.name}: rel L-inf {err}")
        assert err.max() < 1e-5, (w.name, err)


@pytest.mark.parametrize("S", [5, 7])
def test_c3_openmp_compatible_factors(S):
    """An ensemble domain with S = 5 or 7, where the OpenMP backend's y/z
    sweeps read some columns' factors from another substrate, stepped in the
    OpenMP-compatible mode against the oracle's same mode, 20 steps."""
    k = next(i for i in range(64) if wl.c3_domain(i, with_field=False).S == S)
    w = wl.c3_domain(k)
    want, _, _ = _oracle_run(w, 20, quirks=True)
    d = wl.make_domain(w, openmp_quirks=True)
    d.step_graph(20, 10)
    err = rel_linf(d.field(), want)
    _say(f"c3 quirks {w.name}: rel L-inf {err}")
    assert err.max() < 1e-5, err
    # the two modes really differ on this domain
    plain, _, _ = _oracle_run(w, 20, quirks=False)
    assert np.abs(plain - want).max() > 1e-3 * np.abs(want).max()


# ------------------------------------------------------------------ served stream

def test_served_stream_real_sizes():
    """Requests at the served workload's sizes (cube sides 32-256, S 1-6,
    0-200k cells; step counts capped at 40): every result equals the request
    stepped alone bit for bit, and two requests match the oracle."""
    rng = np.random.default_rng(146)
    reqs = serve.poisson_stream(rng, 8, 50.0, step_range=(10, 40))
    _say("served: sides " + ", ".join(f"{r.grid[0]}x{r.S}" for r in reqs))
    srv = serve.Server(chunk=10)
    for r in reqs:
        srv.submit(r)
    done = srv.run()
    srv.close()
    assert sorted(r.rid for r in done) == list(range(len(reqs))), [(r.rid, r.status, r.error) for r in srv.done]
    for r in reqs:
        d = serve.domain_for(r)
        d.step(r.steps)
        assert np.array_equal(d.field(), r.result), r.rid
        d.close()
    big = max(reqs, key=lambda r: r.voxels * r.S)
    other = next(r for r in reqs if r is not big and r.agents is not None)
    for r in (big, other):
        w = serve.workload_of(r)
        want, _, _ = _oracle_run(w, r.steps)
        err = rel_linf(r.result, want)
        _say(f"served request {r.rid} ({r.grid[0]}^3x{r.S}, {r.steps} steps): rel L-inf {err}")
        assert err.max() < 1e-5, err


# ------------------------------------------------------------------ reference benchmark

def test_reference_benchmark_workload():
    """The reference's own benchmark workload: 250^3 x 4, bulk in every voxel,
    Dirichlet on all six faces, 2e6 lattice agents plus one sharing the first
    agent's voxel, internalized tracking on, positions recomputed every 10
    steps; 20 steps against the oracle."""
    w = wl.reference_benchmark()
    steps = 20
    d = wl.make_domain(w)
    for k in range(0, steps, w.recompute_every):
        if k:
            d.recompute_positional_data()
        d.step_graph(w.recompute_every, w.recompute_every)
    want, s, _ = _oracle_run(w, steps, recompute_every=w.recompute_every)
    err = rel_linf(d.field(), want)
    _say(f"reference benchmark: rel L-inf {err}")
    assert err.max() < 2e-6, err
    assert np.array_equal(d.internalized(), s.agents.internalized)  # zero rates: stays 0
    m_gpu, m_or = d.total_mass(), _mass(want, w)
    assert np.all(np.abs(m_gpu - m_or) <= 1e-7 * np.abs(m_or)), (m_gpu, m_or)
