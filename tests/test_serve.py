"""Served form: continuous batching over a request stream.  Each request's
result is the same alone and in the mix; admission respects the HBM budget."""
import numpy as np
import pytest

from physicore_amd import serve


def _small_stream(seed, count=8, rate=200.0):
    rng = np.random.default_rng(seed)
    return serve.poisson_stream(rng, count, rate, side_range=(6, 40), s_range=(1, 6), max_cells=300,
                                step_range=(5, 60))


def test_request_bytes_scale_with_size():
    small = _small_stream(1, count=1)[0]
    big = serve.Request(**{**small.__dict__, "grid": (64, 64, 64), "field": None, "agents": None})
    assert big.device_bytes() > small.device_bytes() > 4 * small.voxels * small.S


def test_submit_rejects_request_over_budget():
    r = _small_stream(2, count=1)[0]
    s = serve.Server(hbm_budget=r.device_bytes() - 1)
    with pytest.raises(ValueError):
        s.submit(r)


@pytest.mark.gpu
def test_results_same_alone_and_in_the_mix():
    from gpu_helpers import require_gpu
    from physicore_amd import bfvm

    require_gpu()
    reqs = _small_stream(3, count=10)
    srv = serve.Server(chunk=4)
    for r in reqs:
        srv.submit(r)
    done = srv.run()
    srv.close()
    assert sorted(r.rid for r in done) == list(range(len(reqs))), [(r.rid, r.status, r.error) for r in srv.done]
    # requests overlapped: some were admitted before others finished
    assert any(a.admitted < b.finished and b.admitted < a.finished
               for a in done for b in done if a.rid < b.rid)
    for r in reqs:
        d = bfvm.Domain(r.dims, r.grid, r.voxel, r.dt, r.diffusion, r.decay, initial=r.initial,
                        reduce=r.reduce)
        if r.field is not None:
            d.upload_field(r.field)
        d.set_dirichlet_faces(*r.faces)
        if r.agents is not None:
            d.set_agents(**r.agents)
        d.step(r.steps)
        assert np.array_equal(d.field(), r.result), r.rid
        assert r.finished >= r.admitted >= 0


@pytest.mark.gpu
def test_admission_respects_budget():
    from gpu_helpers import require_gpu

    require_gpu()
    reqs = _small_stream(4, count=8, rate=1e6)  # all arrive at once
    budget = 2.2 * max(r.device_bytes() for r in reqs)
    srv = serve.Server(hbm_budget=budget, chunk=5)
    for r in reqs:
        srv.submit(r)
    peak = 0.0
    while srv.step_once():
        peak = max(peak, srv.in_use)
        assert srv.in_use <= budget
    assert len(srv.done) == len(reqs)
    assert peak > 0
    ttc = [r.finished - r.arrival for r in srv.done]
    assert np.isfinite(serve.percentile(ttc, 99))


# ------------------------------------------------------- host logic on the CPU
# The loop's bookkeeping (admission, pipelining one chunk ahead, cancel,
# failure isolation, result collection) runs against stand-in domains whose
# "field" counts the steps applied to it, so every request's result says how
# many steps it received.

class _FakeDomain:
    def __init__(self, r, fail_after=None):
        self.shape = (r.grid[2], r.grid[1], r.grid[0], r.S)
        self.steps = 0
        self.fail_after = fail_after
        self.closed = False

    def failing(self):
        return self.fail_after is not None and self.steps > self.fail_after

    def sync(self):
        assert not self.closed
        if self.failing():
            raise serve.bfvm.BfvmError(-9, "agent outside the mesh")

    def field(self):
        self.sync()
        return np.full(self.shape, self.steps, dtype=np.float32)

    def close(self):
        self.closed = True


class _FakeBatch:
    log = []
    shapes = []
    wait_reports = True  # a chunk's wait reports its domains' errors

    def __init__(self, doms):
        assert doms and not any(d.closed for d in doms)
        self.doms = list(doms)
        self.ticks = 0
        self.queried = set()
        self.closed = False

    def step(self, k, graph_steps=0):
        assert not self.closed and not any(d.closed for d in self.doms)
        for d in self.doms:
            d.steps += k
        _FakeBatch.log.append(len(self.doms))
        _FakeBatch.shapes.append({d.shape for d in self.doms})

    def mark(self):
        self.ticks += 1
        return self.ticks - 1

    def wait(self, ticket):
        assert ticket < self.ticks
        if _FakeBatch.wait_reports and any(d.failing() for d in self.doms if not d.closed):
            raise serve.bfvm.BfvmError(-1, "batch failed")

    def query(self, ticket):
        # the latest chunk is still running at its first query
        self.wait(ticket)
        if ticket == self.ticks - 1 and ticket not in self.queried:
            self.queried.add(ticket)
            return False
        return True

    def close(self):
        self.closed = True


@pytest.fixture
def fake_device(monkeypatch):
    import types

    from physicore_amd import bfvm

    made = {}
    fail = {}

    def create(r):
        made[r.rid] = _FakeDomain(r, fail.get(r.rid))
        return made[r.rid]

    monkeypatch.setattr(serve, "bfvm", types.SimpleNamespace(Batch=_FakeBatch,
                                                             BfvmError=bfvm.BfvmError))
    monkeypatch.setattr(serve, "domain_for", create)
    _FakeBatch.log = []
    _FakeBatch.shapes = []
    _FakeBatch.wait_reports = True
    return made, fail


def _tiny(rid, steps, arrival=0.0, S=2, side=4):
    return serve.Request(dims=3, grid=(side, side, side), voxel=(10.0, 10.0, 10.0), dt=0.01,
                         diffusion=np.ones(S), decay=np.full(S, 0.1), steps=steps, arrival=arrival, rid=rid)


def test_every_request_gets_exactly_its_steps(fake_device):
    made, _ = fake_device
    srv = serve.Server(chunk=4)
    steps = [3, 9, 1, 17, 4, 8, 0, 12]
    for i, n in enumerate(steps):
        srv.submit(_tiny(i, n))
    done = srv.run()
    srv.close()
    assert sorted(r.rid for r in done) == list(range(len(steps)))
    for r in done:
        assert r.status == serve.DONE and r.error is None
        assert np.all(r.result == steps[r.rid]), r.rid
    assert all(d.closed for d in made.values())
    assert srv.in_use == 0
    # the batch shrank as requests left, and never ran a request past its count
    assert max(_FakeBatch.log) == len(steps) - 1  # the 0-step request never joins a chunk


def test_malformed_requests_fail_alone():
    good = _tiny(0, 5)
    bad = {
        1: dict(diffusion=np.ones(0), decay=np.ones(0)),
        2: dict(dt=0.0),
        3: dict(dt=float("nan")),
        4: dict(decay=np.array([0.1, np.nan])),
        5: dict(field=np.full((4, 4, 4, 2), np.inf)),
        6: dict(field=np.zeros((3, 4, 4, 2))),
        7: dict(diffusion=np.array([1.0, -1.0])),
        8: dict(grid=(4, 0, 4)),
    }
    srv = serve.Server()
    srv.submit(good)
    for rid, change in bad.items():
        r = _tiny(rid, 5)
        for k, v in change.items():
            setattr(r, k, v)
        srv.submit(r)
    assert [r.rid for r in srv.waiting] == [0]
    assert sorted(r.rid for r in srv.done) == sorted(bad)
    for r in srv.done:
        assert r.status == serve.FAILED and r.error, r.rid
    srv.close()


def test_agent_outside_the_mesh_fails_at_submit():
    r = _tiny(0, 5)
    n = 3
    pos = np.full((n, 3), 5.0)
    pos[1, 2] = 40.0  # the mesh spans [0, 40) micrometres
    r.agents = dict(positions=pos, secretion=np.ones((n, 2)), saturation=np.ones((n, 2)),
                    uptake=np.ones((n, 2)), net_export=np.zeros((n, 2)), internalized=None,
                    fraction_released=np.zeros((n, 2)), volumes=np.ones(n), is_active=np.ones(n, dtype=np.uint8))
    assert serve.validate(r) is not None
    r.agents["is_active"][1] = 0  # an inactive agent may lie anywhere
    assert serve.validate(r) is None


def test_cancel_waiting_and_running_requests(fake_device):
    made, _ = fake_device
    srv = serve.Server(chunk=2)
    for i in range(4):
        srv.submit(_tiny(i, 20, arrival=0.0 if i < 3 else 1e9))
    assert srv.cancel(3)  # still waiting (arrives far in the future)
    assert srv.step_once()
    assert srv.cancel(1)  # running
    done = srv.run()
    srv.close()
    assert sorted(r.rid for r in done) == [0, 2]
    status = {r.rid: r.status for r in srv.done}
    assert status[1] == serve.CANCELLED and status[3] == serve.CANCELLED
    assert 3 not in made and made[1].closed and made[1].steps < 20
    assert not srv.cancel(0)  # already left
    for r in done:
        assert np.all(r.result == 20)


@pytest.mark.parametrize("where", ["chunk wait", "result sync"])
def test_device_failure_stays_with_its_request(fake_device, where):
    made, fail = fake_device
    fail[2] = 5  # request 2's domain reports an error after 5 steps
    _FakeBatch.wait_reports = where == "chunk wait"
    srv = serve.Server(chunk=3)
    for i in range(5):
        srv.submit(_tiny(i, 12))
    done = srv.run()
    srv.close()
    assert sorted(r.rid for r in done) == [0, 1, 3, 4]
    bad = [r for r in srv.done if r.rid == 2][0]
    assert bad.status == serve.FAILED and "outside" in bad.error
    for r in done:
        assert np.all(r.result == 12)
    assert srv.in_use == 0 and all(d.closed for d in made.values())


def test_requests_joining_later_overlap_running_ones(fake_device):
    srv = serve.Server(chunk=2, clock=iter(np.arange(0.0, 1e6, 0.01)).__next__)
    for i in range(6):
        srv.submit(_tiny(i, 10, arrival=0.05 * i))
    done = srv.run()
    srv.close()
    assert len(done) == 6
    assert any(a.admitted < b.finished and b.admitted < a.finished
               for a in done for b in done if a.rid < b.rid)
    assert max(_FakeBatch.log) > 1


@pytest.mark.gpu
def test_bad_request_on_the_device_fails_alone(monkeypatch):
    """A request whose agent lies outside its mesh (let past the submit check)
    fails with the device's error; the requests batched with it finish with
    the results they get alone."""
    from gpu_helpers import require_gpu

    require_gpu()
    reqs = _small_stream(5, count=5, rate=1e6)
    bad = [r for r in reqs if r.agents is not None][0]
    bad.agents = {k: (None if v is None else np.array(v, copy=True)) for k, v in bad.agents.items()}
    bad.agents["positions"][0, 0] = bad.grid[0] * bad.voxel[0] + 50.0
    monkeypatch.setattr(serve, "validate", lambda r: None)
    srv = serve.Server(chunk=4)
    for r in reqs:
        srv.submit(r)
    done = srv.run()
    srv.close()
    assert sorted(r.rid for r in done) == sorted(r.rid for r in reqs if r is not bad)
    assert bad.status == serve.FAILED and bad.error
    for r in done:
        d = serve.domain_for(r)
        d.step(r.steps)
        assert np.array_equal(d.field(), r.result), r.rid
        d.close()


@pytest.mark.gpu
def test_cancel_running_request_on_the_device():
    from gpu_helpers import require_gpu

    require_gpu()
    reqs = _small_stream(6, count=4, rate=1e6)
    for r in reqs:
        r.steps = 40
    srv = serve.Server(chunk=4)
    for r in reqs:
        srv.submit(r)
    while not srv.active:
        assert srv.step_once()
    victim = srv.active[0].req.rid
    assert srv.cancel(victim)
    done = srv.run()
    srv.close()
    assert sorted(r.rid for r in done) == sorted(r.rid for r in reqs if r.rid != victim)
    assert [r.status for r in srv.done if r.rid == victim] == [serve.CANCELLED]
    assert srv.in_use == 0
    for r in done:
        d = serve.domain_for(r)
        d.step(r.steps)
        assert np.array_equal(d.field(), r.result), r.rid
        d.close()


def test_size_classes_step_apart(fake_device):
    """Small and large requests go to different batches (their own streams),
    and each still gets exactly its steps."""
    srv = serve.Server(chunk=3, tiers=(1000,))
    sides = [4, 12, 4, 12, 4, 4, 12]
    for i, side in enumerate(sides):
        srv.submit(_tiny(i, 7 + 2 * i, side=side))
    assert [srv.tier_of(r) for r in srv.waiting] == [0 if s == 4 else 1 for s in sides]
    done = srv.run()
    srv.close()
    assert len(done) == len(sides)
    for r in done:
        assert np.all(r.result == 7 + 2 * r.rid)
    assert all(len(shapes) == 1 for shapes in _FakeBatch.shapes)
    assert {next(iter(sh)) for sh in _FakeBatch.shapes} == {(4, 4, 4, 2), (12, 12, 12, 2)}


def test_batch_follows_membership_when_objects_are_recycled(fake_device, monkeypatch):
    """Domains that have left are dropped, so new ones may reuse their object
    ids: the batch must still be rebuilt whenever its members change (a stale
    batch would step closed domains, which the stand-in batch refuses)."""
    monkeypatch.setattr(serve, "domain_for", lambda r: _FakeDomain(r))
    reqs = [_tiny(i, 2 + i % 3) for i in range(40)]
    srv = serve.Server(chunk=2, tiers=(), hbm_budget=1.5 * reqs[0].device_bytes())  # one at a time
    for r in reqs:
        srv.submit(r)
    done = srv.run()
    srv.close()
    assert len(done) == 40
    for r in done:
        assert np.all(r.result == 2 + r.rid % 3)
