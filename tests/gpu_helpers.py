"""Builds GPU domains (physicore_amd.bfvm.Domain) from the oracle-side scenario
objects so both sides start from identical inputs."""
import numpy as np
import pytest

from physicore_amd import bfvm


def require_gpu():
    n = bfvm.device_count()  # raises if libbfvm_hip.so is missing: no silent skip
    if n < 1:
        pytest.fail("no HIP device visible to a test marked gpu")


def faces_arrays(spec, S):
    vals = np.zeros((6, S))
    conds = np.zeros((6, S), dtype=np.uint8)
    for a in range(3):
        for side, (v, c) in enumerate(((spec.min_values[a], spec.min_conds[a]),
                                       (spec.max_values[a], spec.max_conds[a]))):
            if v is not None and c is not None:
                vals[2 * a + side] = v
                conds[2 * a + side] = c
    return vals, conds


def gpu_domain(p, field=None, dirichlet=None, bulk=None, agents=None, internalized=False,
               precision="fp64", reduce="atomic"):
    d = bfvm.Domain(p.dims, (p.nx, p.ny, p.nz), p.voxel, p.dt, p.diffusion, p.decay,
                    np.zeros(p.S), bbox_min=p.bbox_min, compute_internalized=internalized,
                    precision=precision, reduce=reduce, openmp_quirks=p.openmp_quirks)
    if field is not None:
        d.upload_field(np.asarray(field, dtype=np.float64))
    if dirichlet is not None:
        d.set_dirichlet_faces(*faces_arrays(dirichlet, p.S))
        if len(dirichlet.interior_voxels):
            d.set_dirichlet_interior(dirichlet.interior_voxels, dirichlet.interior_values,
                                     dirichlet.interior_conds)
    if bulk is not None:
        supply, uptake, target = (np.asarray(a, dtype=np.float64).reshape(-1, p.S) for a in bulk)
        active = np.nonzero((supply != 0).any(axis=1) | (uptake != 0).any(axis=1))[0]
        d.set_bulk(active, supply[active], uptake[active], target[active])
    if agents is not None:
        set_agents(d, agents)
    return d


def set_agents(d, ag):
    d.set_agents(ag.positions, ag.secretion, ag.saturation, ag.uptake, ag.net_export,
                 ag.internalized, ag.fraction_released, ag.volumes, ag.is_active)


def rel_linf(got, want):
    """Relative L-infinity error per substrate: max|got-want| / max|want|."""
    got = np.asarray(got, dtype=np.float64).reshape(-1, want.shape[-1])
    want = np.asarray(want, dtype=np.float64).reshape(-1, want.shape[-1])
    scale = np.maximum(np.abs(want).max(axis=0), 1e-300)
    return np.abs(got - want).max(axis=0) / scale
