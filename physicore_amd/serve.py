"""Served form of the BioFVM backend: a request loop with continuous batching.

Requests (one microenvironment each: grid, substrates, coefficients, initial
field, hooks, agents and a step count) arrive over time.  The server admits a
request as soon as its device state fits the HBM budget, keeps that state
resident for the request's lifetime, and steps every admitted request together
in one batch (physicore_amd.bfvm.Batch: one launch per kernel group per sweep),
a chunk of steps at a time.  Requests that arrive while others run join at the
next chunk boundary, so their work overlaps on the device; finished requests
leave, with their result field and time-to-completion.

Each domain's arithmetic in a batch is the arithmetic of stepping it alone, so
a request's result does not depend on what it was batched with (with the
sorted cell reduction the result is bit-identical).
"""
import dataclasses
import time
from typing import Optional

import numpy as np

from physicore_amd import bfvm


@dataclasses.dataclass
class Request:
    """One simulation request.  Arrays follow physicore_amd.bfvm.Domain."""
    dims: int
    grid: tuple
    voxel: tuple
    dt: float
    diffusion: np.ndarray
    decay: np.ndarray
    steps: int
    field: Optional[np.ndarray] = None        # (nz, ny, nx, S), or None: `initial`
    initial: Optional[np.ndarray] = None      # per-substrate initial value (zeros if None)
    faces: Optional[tuple] = None             # (values[6, S], conds[6, S])
    bulk: Optional[tuple] = None              # (voxels, supply, uptake, target)
    agents: Optional[dict] = None             # keyword arguments of Domain.set_agents
    reduce: str = "sorted"
    arrival: float = 0.0                      # seconds, on the server clock
    rid: int = -1
    # filled by the server
    admitted: float = float("nan")
    finished: float = float("nan")
    result: Optional[np.ndarray] = None

    @property
    def S(self):
        return len(self.diffusion)

    @property
    def voxels(self):
        return int(np.prod(self.grid))

    def device_bytes(self):
        """Upper estimate of the device memory the request holds: the FP32
        field, the FP64 factor tables, the agent arrays with their FP64 per-cell
        and per-owner sums, the voxel ballot and the sort buffers."""
        V, S = self.voxels, self.S
        n = 0 if self.agents is None else len(self.agents["volumes"])
        b = 4 * V * S + 4 * V                     # field, ballot
        b += 3 * 3 * 8 * max(self.grid) * S         # b', e, c per axis
        b += n * (8 * self.dims + 8 + 1)            # positions, volumes, activity
        b += n * S * 8 * (7 + 10)                   # rates/internalized + sums
        b += n * 4 * 12                             # voxel, owner, keys, scans
        if self.bulk is not None:
            b += len(self.bulk[0]) * (8 + 16 * S)
        return b + (1 << 20)


class Server:
    """Continuous-batching request loop over one GPU."""

    def __init__(self, hbm_budget=200e9, chunk=10, graph=True, clock=time.perf_counter):
        self.budget = float(hbm_budget)
        self.chunk = int(chunk)
        self.graph = bool(graph)
        self.clock = clock
        self.t0 = clock()
        self.waiting = []          # submitted, not admitted (arrival order)
        self.active = []           # (request, domain, steps done)
        self.done = []
        self.in_use = 0.0
        self.updates = 0           # voxel*substrate updates performed
        self._batch = None
        self._batch_key = None
        self.chunks = 0

    def now(self):
        return self.clock() - self.t0

    def submit(self, req: Request):
        if req.device_bytes() > self.budget:
            raise ValueError(f"request {req.rid} needs {req.device_bytes()} bytes, over the budget")
        self.waiting.append(req)

    def _admit(self):
        t = self.now()
        still = []
        for r in self.waiting:
            if r.arrival <= t and self.in_use + r.device_bytes() <= self.budget:
                d = domain_for(r)
                r.admitted = self.now()
                self.in_use += r.device_bytes()
                self.active.append([r, d, 0])
            else:
                still.append(r)
        self.waiting = still

    def _retire(self):
        keep = []
        for item in self.active:
            r, d, k = item
            if k >= r.steps:
                r.result = d.field()
                r.finished = self.now()
                self.in_use -= r.device_bytes()
                self.done.append(r)
            else:
                keep.append(item)
        self.active = keep

    def step_once(self):
        """Admits what has arrived and fits, runs one chunk of the batch and
        retires finished requests.  Returns False when nothing is left."""
        self._admit()
        if not self.active:
            if not self.waiting:
                return False
            # idle until the next arrival
            nxt = min(r.arrival for r in self.waiting)
            delay = nxt - self.now()
            if delay > 0:
                time.sleep(min(delay, 0.05))
            return True
        # a chunk never runs a request past its step count
        k = min(self.chunk, min(r.steps - done for r, _, done in self.active))
        key = tuple(id(d) for _, d, _ in self.active)
        if key != self._batch_key:
            if self._batch is not None:
                self._batch.close()
            self._batch = bfvm.Batch([d for _, d, _ in self.active])
            self._batch_key = key
        self._batch.step(k, graph_steps=self.chunk if (self.graph and k == self.chunk) else 0)
        self._batch.sync()
        self.chunks += 1
        for item in self.active:
            item[2] += k
            self.updates += k * item[0].voxels * item[0].S
        self._retire()
        return True

    def run(self):
        while self.step_once():
            pass
        if self._batch is not None:
            self._batch.close()
            self._batch = None
        return self.done


def domain_for(r):
    """A bfvm.Domain holding request r's microenvironment (field, hooks, agents)."""
    d = bfvm.Domain(r.dims, r.grid, r.voxel, r.dt, r.diffusion, r.decay, initial=r.initial, reduce=r.reduce)
    if r.field is not None:
        d.upload_field(r.field)
    if r.faces is not None:
        d.set_dirichlet_faces(*r.faces)
    if r.bulk is not None:
        d.set_bulk(*r.bulk)
    if r.agents is not None:
        d.set_agents(**r.agents)
    return d


def workload_of(r):
    """Request r as a physicore_amd.workloads.Workload (its step count included)."""
    from physicore_amd import workloads

    return workloads.Workload(f"request {r.rid}", r.dims, tuple(r.grid), tuple(r.voxel), r.dt,
                              np.asarray(r.diffusion), np.asarray(r.decay), r.steps,
                              initial=None if r.initial is None else np.asarray(r.initial), field=r.field,
                              faces=r.faces, bulk=r.bulk, agents=r.agents)


def percentile(values, q):
    return float(np.percentile(np.asarray(values, dtype=np.float64), q)) if len(values) else float("nan")


def poisson_stream(rng, count, rate, side_range=(32, 256), s_range=(1, 6), max_cells=200_000,
                   step_range=(100, 5000), dx=20, dt=0.01, random_field=True):
    """Requests of the served workload: cube side log-uniform in side_range,
    S uniform in s_range, D log-uniform in [1, 1e5], decay uniform in
    [0.01, 1], field U[0,1] (random_field) or a per-substrate constant drawn
    U[0,1], Dirichlet 1.0 on x-min, 0..max_cells uniform cells
    (secretion/uptake U[0,10], saturation 1, volume 2494), steps log-uniform
    in step_range; exponential inter-arrival times at `rate` per second."""
    reqs = []
    t = 0.0
    for i in range(count):
        t += rng.exponential(1.0 / rate)
        side = int(round(np.exp(rng.uniform(np.log(side_range[0]), np.log(side_range[1])))))
        S = int(rng.integers(s_range[0], s_range[1] + 1))
        D = 10.0 ** rng.uniform(0.0, 5.0, S)
        lam = rng.uniform(0.01, 1.0, S)
        steps = int(round(np.exp(rng.uniform(np.log(step_range[0]), np.log(step_range[1])))))
        vals = np.zeros((6, S))
        conds = np.zeros((6, S), dtype=np.uint8)
        vals[0], conds[0] = 1.0, 1
        n = int(rng.integers(0, max_cells + 1))
        box = side * dx
        agents = None
        if n:
            pos = np.minimum(rng.uniform(0.0, box, (n, 3)), np.nextafter(box, 0))
            agents = dict(positions=pos, secretion=rng.uniform(0, 10, (n, S)), saturation=np.ones((n, S)),
                          uptake=rng.uniform(0, 10, (n, S)), net_export=np.zeros((n, S)), internalized=None,
                          fraction_released=np.zeros((n, S)), volumes=np.full(n, 2494.0),
                          is_active=np.ones(n, dtype=np.uint8))
        field = rng.uniform(0.0, 1.0, (side, side, side, S)) if random_field else None
        initial = None if random_field else rng.uniform(0.0, 1.0, S)
        reqs.append(Request(dims=3, grid=(side, side, side), voxel=(dx, dx, dx), dt=dt, diffusion=D, decay=lam,
                            steps=steps, field=field, initial=initial, faces=(vals, conds), agents=agents,
                            arrival=t, rid=i))
    return reqs
