"""Field snapshots for output (bfvm_snapshot_begin/wait): a snapshot holds the
field of the step it was queued after, while the steps queued behind it run
during its copy to pinned host memory."""
import numpy as np
import pytest

import scenarios as sc
from gpu_helpers import gpu_domain, require_gpu

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _gpu():
    require_gpu()


def test_snapshots_hold_the_step_they_were_taken_at():
    from physicore_amd import bfvm

    p, f0, dirs, bulk, ag = sc.config2(side=64, S=3, n_cells=2000, seed=21)
    # the sorted reduction makes the two domains' cell sums bit-identical
    a = gpu_domain(p, field=f0, dirichlet=dirs, bulk=bulk, agents=ag, reduce="sorted")
    b = gpu_domain(p, field=f0, dirichlet=dirs, bulk=bulk, agents=ag, reduce="sorted")
    buf0, buf1 = bfvm.PinnedField(a.shape), bfvm.PinnedField(a.shape)
    a.step(5)
    t0 = a.snapshot_begin(buf0)
    a.step_graph(20, 10)  # queued behind the snapshot
    t1 = a.snapshot_begin(buf1)  # waits on the device for the first host copy
    a.step(3)
    a.snapshot_wait(t1)
    a.snapshot_wait(t0)
    b.step(5)
    assert np.array_equal(buf0.array, b.field())
    b.step_graph(20, 10)
    assert np.array_equal(buf1.array, b.field())
    b.step(3)
    assert np.array_equal(a.field(), b.field())
    with pytest.raises(bfvm.BfvmError):
        a.snapshot_wait(t1 + 1)  # not begun
    buf0.close()
    buf1.close()


def test_snapshot_after_upload_and_shape_check():
    from physicore_amd import bfvm

    p, f0, dirs, bulk, _ = sc.config2(side=24, S=2, n_cells=0, seed=22)
    d = gpu_domain(p, field=f0, dirichlet=dirs, bulk=bulk)
    want = np.random.default_rng(3).uniform(0, 1, d.shape).astype(np.float32)
    d.upload_field(want)
    buf = bfvm.PinnedField(d.shape)
    d.snapshot_wait(d.snapshot_begin(buf))
    assert np.array_equal(buf.array, want)
    with pytest.raises(ValueError):
        d.snapshot_begin(bfvm.PinnedField((1, 2, 3, 4)))
    buf.close()
