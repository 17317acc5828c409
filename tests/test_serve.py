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
    assert sorted(r.rid for r in done) == list(range(len(reqs)))
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
