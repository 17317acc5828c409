"""bfvm_total_mass_static: the mass reduction with its scratch in the library's
static device array instead of allocated buffers, against bfvm_total_mass."""
import numpy as np
import pytest

import oracle as orc
from gpu_helpers import gpu_domain, require_gpu
from physicore_amd import bfvm

pytestmark = pytest.mark.gpu


def _domain(grid, S, seed):
    rng = np.random.default_rng(seed)
    p = orc.Problem(3, grid, (20, 20, 20), 0.01, 10 ** rng.uniform(0, 5, S), rng.uniform(0.01, 1, S))
    return gpu_domain(p, field=rng.uniform(0, 1, p.shape))


def test_static_mass_equals_allocating_mass_twice():
    """Two calls in a row, the second on a smaller domain (fewer partial sums
    than the first call left in the shared array): both equal bfvm_total_mass
    bit for bit, before and after steps."""
    require_gpu()
    big = _domain((200, 150, 100), 6, 1)     # 46 blocks x 6 substrates of partials
    small = _domain((20, 17, 9), 3, 2)       # 1 block x 3 substrates
    for d in (big, small):
        a = d.total_mass_static()
        b = d.total_mass()
        assert np.array_equal(a, b), (a, b)
        d.step(2)
        assert np.array_equal(d.total_mass_static(), d.total_mass())
    # the small domain's call after the big one again, then the big one
    assert np.array_equal(small.total_mass_static(), small.total_mass())
    assert np.array_equal(big.total_mass_static(), big.total_mass())


def test_static_mass_rejects_pending_releases():
    require_gpu()
    p = orc.Problem(1, (5, 1, 1), (20, 20, 20), 0.01, [0.0], [0.0])
    ag = orc.AgentSpec(1, 1, 1)
    ag.positions[0] = [30.0]
    ag.volumes[0] = 1000.0
    ag.internalized[0] = 8.0
    ag.fraction_released[0] = 1.0
    d = gpu_domain(p, field=np.ones(p.shape), agents=ag, internalized=True)
    d.step(1)
    d.release_internalized(0)
    with pytest.raises(bfvm.BfvmError) as e:
        d.total_mass_static()
    assert "releases are pending" in str(e.value)
    m = d.total_mass()  # applies the release
    assert np.array_equal(d.total_mass_static(), m)
