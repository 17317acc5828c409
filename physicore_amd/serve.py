"""Served form of the BioFVM backend: a request loop with continuous batching.

Requests (one microenvironment each: grid, substrates, coefficients, initial
field, hooks, agents and a step count) arrive over time.  The server admits a
request as soon as its device state fits the HBM budget, keeps that state
resident for the request's lifetime, and steps the admitted requests together
in batches (physicore_amd.bfvm.Batch: one launch per kernel group per sweep),
a chunk of steps at a time.  Requests that arrive while others run join at the
next chunk boundary, so their work overlaps on the device; finished requests
leave with their result field and time-to-completion.

The host stays ahead of the GPU: it queues chunk i while chunk i-1 runs
(batch tickets), so the GPU never waits for the host between chunks
(optionally per size class, each class a batch on its own streams).  New
requests are created and uploaded while chunks run (their own streams), and
results are copied out while later chunks run.

Each domain's arithmetic in a batch is the arithmetic of stepping it alone, so
a request's result does not depend on what it was batched with (with the
sorted cell reduction the result is bit-identical).

Failures stay with their request: a malformed request (no substrates, a
non-positive or non-finite time step, non-finite coefficients, fields or
rates, an active agent outside its mesh) fails at submit; an error the device
reports for one request fails that request only; a non-finite result fails
its request.  cancel() withdraws a waiting or running request.
"""
import dataclasses
import time
from typing import Optional

import numpy as np

from physicore_amd import bfvm

WAITING, ACTIVE, DONE, FAILED, CANCELLED = "waiting", "active", "done", "failed", "cancelled"


@dataclasses.dataclass(eq=False)
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
    status: str = WAITING
    error: Optional[str] = None
    admitted: float = float("nan")
    finished: float = float("nan")
    result: Optional[np.ndarray] = None
    # server clock at each stage: "uploaded" (joined a batch), "queued"
    # (last chunk queued), "copied" (result on the host)
    times: dict = dataclasses.field(default_factory=dict)

    @property
    def S(self):
        return len(self.diffusion)

    @property
    def voxels(self):
        return int(np.prod(self.grid))

    def device_bytes(self):
        """Upper estimate of the device memory the request holds: the FP32
        field (a reused block may be up to a quarter larger), the FP64 factor
        tables, the agent arrays with their FP64 per-cell and per-owner sums,
        the voxel ballot and the sort buffers."""
        V, S = self.voxels, self.S
        n = 0 if self.agents is None else len(self.agents["volumes"])
        b = 5 * V * S + 4 * V                     # field, ballot
        b += 3 * 3 * 8 * max(self.grid) * S         # b', e, c per axis
        b += n * (8 * self.dims + 8 + 1)            # positions, volumes, activity
        b += n * S * 8 * (7 + 10)                   # rates/internalized + sums
        b += n * 4 * 12                             # voxel, owner, keys, scans
        if self.bulk is not None:
            b += len(self.bulk[0]) * (8 + 16 * S)
        return b + (1 << 20)


def _finite(*arrays):
    return all(a is None or np.all(np.isfinite(np.asarray(a, dtype=np.float64))) for a in arrays)


def validate(r):
    """Why request r cannot run (None if it can): the checks bfvm_create and
    the agent binning would fail on, made before any device work, so that a
    bad request fails alone."""
    try:
        S = r.S
        if S < 1:
            return "a request needs at least one substrate"
        if len(r.decay) != S:
            return "diffusion and decay give different substrate counts"
        if not (r.dims in (1, 2, 3)) or len(r.grid) != 3 or any(int(g) < 1 for g in r.grid):
            return "bad mesh"
        if not (np.isfinite(r.dt) and r.dt > 0):
            return "the time step must be positive and finite"
        if not _finite(r.diffusion, r.decay) or np.any(np.asarray(r.diffusion) < 0) or np.any(np.asarray(r.decay) < 0):
            return "coefficients must be finite and non-negative"
        if r.steps < 0:
            return "negative step count"
        if r.field is not None:
            if np.asarray(r.field).size != r.voxels * S:
                return "field size does not match the mesh"
            if not _finite(r.field):
                return "non-finite initial field"
        if r.initial is not None and not _finite(r.initial):
            return "non-finite initial values"
        if r.faces is not None and not _finite(r.faces[0]):
            return "non-finite Dirichlet values"
        if r.bulk is not None and not _finite(*r.bulk[1:]):
            return "non-finite bulk rates"
        a = r.agents
        if a is not None and len(a["volumes"]):
            if not _finite(a["positions"], a["secretion"], a["saturation"], a["uptake"], a["net_export"],
                           a.get("internalized"), a["fraction_released"], a["volumes"]):
                return "non-finite agent data"
            pos = np.asarray(a["positions"], dtype=np.float64).reshape(len(a["volumes"]), -1)[:, : r.dims]
            active = np.asarray(a["is_active"]).astype(bool)
            box = np.asarray(r.grid[: r.dims], dtype=np.float64) * np.asarray(r.voxel[: r.dims], dtype=np.float64)
            q = pos[active] / np.asarray(r.voxel[: r.dims], dtype=np.float64)
            if q.size and (np.any(q < 0) or np.any(q >= np.asarray(r.grid[: r.dims])) or np.any(pos[active] >= box)):
                return "an active agent lies outside the mesh"
    except (TypeError, ValueError, KeyError, AttributeError) as e:
        return f"malformed request: {e}"
    return None


class _Entry:
    """An admitted request: its domain and the steps queued for it.  `serial`
    names the entry for the life of the server (object ids are reused once
    an entry is gone, so a batch is keyed by serials)."""

    _next = 0

    def __init__(self, req, dom):
        self.req = req
        self.dom = dom
        self.done = 0
        self.serial = _Entry._next
        _Entry._next += 1


def _retrieve(server, entry):
    """Helper-thread job: the request's field, copied out after its last
    queued chunk (the context's stream is ordered after it), then the domain
    is freed (its blocks go to the library's cache, no device-wide wait).
    Returns the result array."""
    d = entry.dom
    try:
        res = d.field()
        entry.req.times["copied"] = server.now()
    finally:
        with server._device_lock:
            d.close()
    entry.req.finished = server.now()
    return res


def _create(server, r):
    """Helper-thread job: request r's domain, created and uploaded."""
    with server._device_lock:
        return domain_for(r)


class _Tier:
    """Requests of one size class stepped in lockstep by one batch, a chunk at
    a time, on the batch's own streams."""

    def __init__(self, index):
        self.index = index
        self.active = []       # _Entry: admitted and uploaded, steps still to queue
        self.batch = None
        self.key = None        # the domains the batch was made for
        self.fresh = False     # the batch has not stepped yet
        self.inflight = []     # (batch, ticket) of the queued chunks, oldest first


class Server:
    """Continuous-batching request loop over one GPU.

    Requests can be split by size class (voxel*substrates below each bound in
    `tiers`): each class is a batch stepped in lockstep, and the classes run
    on their own streams side by side, so that a small request need not keep
    the step rate of the largest one it shares the GPU with.  By default there
    is one class: on the served stream (32 Poisson requests, sides 32-256)
    two bounds, 2^20 and 2^24, gave the same p50 and p99 within the spread of
    repeated runs, since the classes' streams share the GPU's four hardware
    queues and their kernels mostly run back to back.  The calling thread
    admits requests, forms the batches and queues chunks, staying `ahead`
    chunks ahead of the GPU in each class.  Two helper threads do the host
    work that would otherwise leave the GPU idle between chunks: one creates
    and uploads admitted requests, the other copies results out and frees
    their device memory (a free may wait for the device).  Domains and
    batches leave through the second one too.  Allocation and freeing on the
    helper threads never overlap the capture of a batch's graph (one lock:
    the capture happens in the first step of a new batch)."""

    def __init__(self, hbm_budget=200e9, chunk=10, graph=True, clock=time.perf_counter,
                 tiers=(), ahead=2):
        import threading
        from concurrent.futures import ThreadPoolExecutor

        self.budget = float(hbm_budget)
        self.chunk = int(chunk)
        self.graph = bool(graph)
        self.clock = clock
        self.bounds = tuple(sorted(tiers))
        self.ahead = max(1, int(ahead))
        self.t0 = clock()
        self.waiting = []          # submitted, not admitted (arrival order)
        self.tiers = [_Tier(k) for k in range(len(self.bounds) + 1)]
        self.done = []             # finished, failed or cancelled requests
        self.in_use = 0.0
        self.updates = 0           # voxel*substrate updates performed
        self.chunks = 0
        self._admitting = []       # (request, future of its Domain)
        self._leaving = []         # (entry, future of its result)
        self._old_batches = []     # replaced batches, freed once their chunks have run
        self._cancel = set()
        self._device_lock = threading.Lock()
        self._admit_pool = ThreadPoolExecutor(max_workers=1, thread_name_prefix="bfvm-admit")
        self._retire_pool = ThreadPoolExecutor(max_workers=1, thread_name_prefix="bfvm-retire")

    def now(self):
        return self.clock() - self.t0

    @property
    def active(self):
        """Admitted requests with steps still to queue (_Entry), all classes."""
        return [e for t in self.tiers for e in t.active]

    def tier_of(self, r):
        u = r.voxels * r.S
        return sum(u >= b for b in self.bounds)

    # ------------------------------------------------------------ requests
    def submit(self, req: Request):
        """Queues a request.  Over the whole budget: ValueError.  Malformed:
        the request fails at once (status FAILED, error set), the others are
        unaffected."""
        if req.device_bytes() > self.budget:
            raise ValueError(f"request {req.rid} needs {req.device_bytes()} bytes, over the budget")
        why = validate(req)
        if why is not None:
            self._finish(req, FAILED, why)
            return
        req.status = WAITING
        self.waiting.append(req)

    def cancel(self, rid):
        """Withdraws request rid: a waiting one at once, an admitted one at the
        next chunk boundary (its memory is freed once its queued work ends).
        Returns False if it has already left."""
        for r in self.waiting:
            if r.rid == rid:
                self.waiting.remove(r)
                self._finish(r, CANCELLED, None)
                return True
        if any(e.req.rid == rid for e in self.active) or any(r.rid == rid for r, _ in self._admitting):
            self._cancel.add(rid)
            return True
        return False

    def _finish(self, r, status, error):
        r.status = status
        r.error = error
        if not np.isfinite(r.finished):
            r.finished = self.now()
        self.done.append(r)

    def _free(self, obj):
        """Closes a domain or batch on the helper thread, after its work."""
        def close():
            with self._device_lock:
                obj.close()
        self._retire_pool.submit(close)

    def _free_old_batches(self):
        busy = {id(b) for t in self.tiers for b, _ in t.inflight}
        keep = []
        for b in self._old_batches:
            if id(b) in busy:
                keep.append(b)
            else:
                self._free(b)
        self._old_batches = keep

    # ---------------------------------------------------------------- loop
    def _admit(self, block=False):
        t = self.now()
        still = []
        for r in self.waiting:
            if r.arrival <= t and self.in_use + r.device_bytes() <= self.budget:
                r.admitted = self.now()
                r.status = ACTIVE
                self.in_use += r.device_bytes()
                self._admitting.append((r, self._admit_pool.submit(_create, self, r)))
            else:
                still.append(r)
        self.waiting = still
        if block and self._admitting:
            self._admitting[0][1].result()  # nothing to run until one is uploaded
        keep = []
        for r, fut in self._admitting:
            if not fut.done():
                keep.append((r, fut))
                continue
            try:
                d = fut.result()
            except Exception as e:  # the device refused this request only
                self.in_use -= r.device_bytes()
                self._finish(r, FAILED, str(e))
                continue
            if r.rid in self._cancel:
                self._cancel.discard(r.rid)
                self.in_use -= r.device_bytes()
                self._finish(r, CANCELLED, None)
                self._free(d)
                continue
            r.times["uploaded"] = self.now()
            self.tiers[self.tier_of(r)].active.append(_Entry(r, d))
        self._admitting = keep

    def _collect(self, block=False):
        keep = []
        for e, fut in self._leaving:
            if not (block or fut.done()):
                keep.append((e, fut))
                continue
            self.in_use -= e.req.device_bytes()
            try:
                res = fut.result()
            except Exception as err:
                self._finish(e.req, FAILED, str(err))
                continue
            if not np.all(np.isfinite(res)):
                self._finish(e.req, FAILED, "non-finite result")
                continue
            e.req.result = res
            self._finish(e.req, DONE, None)
        self._leaving = keep

    def _retire_chunks(self, t, block=False):
        """Drops the class's chunks that have run (with block, waits for the
        oldest first).  A device error fails the requests it came from."""
        while t.inflight:
            batch, ticket = t.inflight[0]
            try:
                if block:
                    batch.wait(ticket)
                    block = False
                elif not batch.query(ticket):
                    return
            except bfvm.BfvmError as err:
                self._isolate_failures(t, err)
                return
            t.inflight.pop(0)

    def _isolate_failures(self, t, batch_error):
        """Fails the class's requests whose domains report an error; if none
        does, the error belongs to the batch and fails all of them.  Syncing
        every domain also waits for every queued chunk (each context's stream
        is ordered after the batch's work), so nothing is in flight after."""
        failed = []
        for e in list(t.active):
            try:
                e.dom.sync()
            except bfvm.BfvmError as err:
                failed.append((e, str(err)))
        if not failed:
            failed = [(e, str(batch_error)) for e in t.active]
        for e, why in failed:
            t.active.remove(e)
            self.in_use -= e.req.device_bytes()
            self._finish(e.req, FAILED, why)
            self._free(e.dom)
        t.key = None
        t.inflight = []

    def _queue_chunk(self, t):
        # a chunk never runs a request past its step count
        k = min(self.chunk, min(e.req.steps - e.done for e in t.active))
        key = tuple(e.serial for e in t.active)
        if key != t.key:
            if t.batch is not None:
                self._old_batches.append(t.batch)
            t.batch = bfvm.Batch([e.dom for e in t.active])
            t.key = key
            t.fresh = True
        batch = t.batch
        graph_steps = self.chunk if (self.graph and k == self.chunk) else 0
        try:
            if t.fresh and graph_steps:
                with self._device_lock:  # captures the batch's graph
                    batch.step(k, graph_steps=graph_steps)
                t.fresh = False
            else:
                batch.step(k, graph_steps=graph_steps)
            ticket = batch.mark()
        except bfvm.BfvmError as err:
            self._isolate_failures(t, err)
            return
        t.inflight.append((batch, ticket))
        self.chunks += 1
        for e in list(t.active):
            e.done += k
            self.updates += k * e.req.voxels * e.req.S
            if e.done >= e.req.steps:
                e.req.times["queued"] = self.now()
                t.active.remove(e)
                self._leaving.append((e, self._retire_pool.submit(_retrieve, self, e)))

    def step_once(self):
        """Admits what has arrived and fits, queues a chunk in every class that
        is fewer than `ahead` chunks ahead of the GPU, drops the chunks that
        have run and collects finished results; when every class is ahead,
        waits for one chunk.  Returns False when nothing is left."""
        busy = any(t.active or t.inflight for t in self.tiers)
        self._admit(block=not busy)
        for t in self.tiers:
            for e in [e for e in t.active if e.req.rid in self._cancel]:
                t.active.remove(e)
                self._cancel.discard(e.req.rid)
                self.in_use -= e.req.device_bytes()
                self._finish(e.req, CANCELLED, None)
                self._free(e.dom)
        self._collect()
        for t in self.tiers:
            self._retire_chunks(t)
        queued = False
        for t in self.tiers:
            if t.active and len(t.inflight) < self.ahead:
                self._queue_chunk(t)
                queued = True
        self._free_old_batches()
        if queued:
            return True
        running = [t for t in self.tiers if t.inflight]
        if not any(t.active for t in self.tiers):
            if running:
                for t in running:
                    self._retire_chunks(t, block=True)
                return True
            if not self.waiting and not self._admitting:
                self._collect(block=True)
                return False
            if not self._admitting:
                # idle until the next arrival
                delay = min(r.arrival for r in self.waiting) - self.now()
                if delay > 0:
                    time.sleep(min(delay, 0.05))
            return True
        # every class with work is `ahead` chunks ahead of the GPU
        if len(running) == 1:
            self._retire_chunks(running[0], block=True)
        else:
            time.sleep(1e-4)
        return True

    def run(self):
        """Serves until every submitted request has left; returns the finished
        ones (failed and cancelled ones are in self.done with their status)."""
        while self.step_once():
            pass
        for t in self.tiers:
            while t.inflight:
                self._retire_chunks(t, block=True)
        self._collect(block=True)
        for t in self.tiers:
            if t.batch is not None:
                self._old_batches.append(t.batch)
            t.batch = None
            t.key = None
        self._free_old_batches()
        self._retire_pool.submit(lambda: None).result()  # the frees queued so far
        return [r for r in self.done if r.status == DONE]

    def close(self):
        self._admit_pool.shutdown(wait=True)
        self._retire_pool.shutdown(wait=True)


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
