"""Parity of the HIP backend with the CPU oracle, through the C ABI.

Tolerances: the field is FP32 in HBM, the oracle FP64.  One step is held to a
relative L-infinity error per substrate of 2e-6 (a few FP32 roundings); multi-step
runs to the 1e-5 bar of the benchmark.  Cell sums are FP64 on both sides: the
sorted reduction must match the oracle bit for bit, the atomic one to 1e-13.
"""
import dataclasses

import numpy as np
import pytest

import oracle as orc
import scenarios as sc
from gpu_helpers import gpu_domain, rel_linf, require_gpu, set_agents

pytestmark = pytest.mark.gpu
G = sc.GOLDEN


@pytest.fixture(autouse=True)
def _gpu():
    require_gpu()


# ------------------------------------------------------------------ diffusion

@pytest.mark.parametrize("name,dims", [("Uniform1D", 1), ("Uniform2D", 2), ("Uniform3D", 3)])
def test_uniform_golden(name, dims):
    t = G["diffusion"][name]
    env = G["diffusion"][t["env"]]
    p = sc.problem_from_env(env, dims, t["bbox_max"], t["voxel"])
    d = gpu_domain(p, field=sc.constant_field(p, env["initial"]))
    d.step(1)
    f = d.field(np.float64)
    for s, want in enumerate(t["expected_per_substrate"]):
        assert np.abs(f[..., s] - want).max() < max(t["tol"], 4e-7 * want)


@pytest.mark.parametrize("name,dims", [("Random1D", 1), ("Random2D", 2), ("Random3D", 3)])
def test_random_golden(name, dims):
    t = G["diffusion"][name]
    env = G["diffusion"][t["env"]]
    p = sc.problem_from_env(env, dims, t["bbox_max"], t["voxel"])
    d = gpu_domain(p, field=sc.linear_index_field(p))
    d.step(1)
    got = d.field(np.float64).ravel()
    want = np.asarray(t["expected"])
    # FP32 storage: 1e-6 absolute, or a few ulps of the value
    assert np.all(np.abs(got - want) <= np.maximum(t["tol"], 2.5e-7 * np.abs(want)))


SHAPES = [
    (3, (5, 4, 3), 1), (3, (7, 9, 11), 2), (3, (33, 17, 9), 3), (3, (64, 64, 64), 4),
    (3, (50, 50, 50), 1), (3, (13, 70, 5), 5), (3, (40, 6, 31), 7), (3, (16, 16, 16), 8),
    (3, (1, 9, 12), 2), (3, (9, 1, 12), 3), (3, (9, 12, 1), 2), (2, (37, 29, 1), 3),
    (1, (101, 1, 1), 2), (3, (129, 3, 4), 1), (3, (3, 130, 5), 2), (3, (4, 5, 200), 4),
    (3, (200, 6, 5), 4), (3, (257, 3, 2), 2), (3, (300, 4, 3), 1), (3, (77, 5, 4), 6),
    (3, (6, 300, 3), 2), (3, (5, 4, 280), 3), (3, (130, 7, 3), 8), (1, (600, 1, 1), 3),
    # y/z lines of 257..512 (two lanes per line): odd and even lengths, both
    # halves' bucket edges, the longest line
    (3, (5, 257, 3), 2), (3, (3, 401, 4), 3), (3, (4, 3, 512), 1), (3, (7, 5, 385), 2),
    (2, (9, 333, 1), 2),
    # y lines of 209..256 in 3-D domains (two lanes per line, half-line
    # buckets 112..128), and the same length as the last (2-D) axis, which
    # keeps one lane per line
    (3, (5, 209, 3), 2), (3, (4, 256, 3), 3), (3, (6, 230, 2), 1), (3, (3, 241, 4), 4),
    (2, (5, 233, 1), 2), (3, (4, 3, 230), 2), (3, (5, 4, 209), 3),
]


# kernel variants: the default (register lines, x kernel per shape), each
# explicit x kernel (strided register lines, LDS transpose, affine scan,
# pipelined, chunked transpose), pipelined y/z lines, round-to-nearest stores,
# and every line solved inside an LDS tile
VARIANTS = {"default": {}, "xstrided": {"x_sweep": 1}, "xrows": {"x_sweep": 2}, "xscan": {"x_sweep": 3},
            "xpipe": {"x_sweep": 4}, "xchunk": {"x_sweep": 5}, "xchunk4": {"x_sweep": 6}, "xtwo": {"x_sweep": 7},
            "pipe": {"pipeline": 1},
            "nearest": {"round_nearest": 1}, "lds": {"force_lds_sweeps": 1}}


@pytest.mark.parametrize("dims,grid,S", SHAPES)
@pytest.mark.parametrize("quirks", [False, True])
@pytest.mark.parametrize("variant", list(VARIANTS))
def test_diffusion_step_matches_oracle(dims, grid, S, quirks, variant):
    rng = np.random.default_rng(hash((dims, grid, S)) % 2**32)
    D = 10 ** rng.uniform(0, 5, S)
    lam = rng.uniform(0.01, 1, S)
    p = orc.Problem(dims, grid, (20, 20, 20), 0.01, D, lam, openmp_quirks=quirks)
    f0 = rng.uniform(0, 1, p.shape)
    want = p.diffuse(f0.astype(np.float32).astype(np.float64))
    d = gpu_domain(p, field=f0)
    if VARIANTS[variant]:
        d.set_tuning(**VARIANTS[variant])
    d.step(1)
    err = rel_linf(d.field(), want)
    assert err.max() < 2e-6, err


# rows for the chunked x kernels: every tail chunk (8, 16, 24, 32 voxels; for
# 16-byte accesses a bucket rounded up past a 24-voxel tail), row counts that
# leave a partial last wave (64/S rows per wave, an even number for odd S),
# exactly one wave, 2-D, the longest rows of each S, and shapes they do not
# take (too few rows, rows past the bucket limit, S = 2) where the strided
# kernel runs instead
CHUNK_SHAPES = [
    (3, (200, 13, 3), 4), (3, (184, 9, 2), 8), (3, (161, 5, 4), 4), (3, (8, 17, 3), 8),
    (3, (40, 16, 1), 4), (2, (97, 23, 1), 4), (3, (59, 7, 3), 8), (3, (128, 8, 1), 8),
    (3, (192, 31, 2), 4), (1, (150, 1, 1), 4), (3, (60, 10, 10), 2),
    (3, (200, 9, 5), 3), (3, (193, 11, 2), 6), (3, (185, 13, 2), 5), (3, (192, 5, 2), 7),
    (3, (200, 5, 2), 7), (2, (24, 20, 1), 3), (3, (8, 3, 3), 5), (3, (100, 7, 9), 3),
    (3, (37, 6, 4), 6), (3, (70, 9, 3), 7), (3, (120, 4, 7), 5), (2, (16, 57, 1), 6),
    (3, (176, 70, 1), 1), (3, (177, 70, 1), 1), (3, (33, 65, 2), 1), (2, (100, 64, 1), 1),
    (3, (200, 33, 2), 2), (3, (17, 40, 1), 2), (2, (129, 31, 1), 2),
]
CHUNK_SHAPES.remove((3, (177, 70, 1), 1))  # rows past one lane: the two-lane kernel (below)
CHUNK_SHAPES.remove((3, (200, 5, 2), 7))


@pytest.mark.parametrize("dims,grid,S", CHUNK_SHAPES)
def test_chunked_x_sweep_matches_strided(dims, grid, S):
    """The chunked x kernel runs the strided kernel's recurrences with its
    factors and rounding keys: fields identical bit for bit over several steps"""
    rng = np.random.default_rng(hash((dims, grid, S, 5)) % 2**32)
    D = 10 ** rng.uniform(0, 5, S)
    lam = rng.uniform(0.01, 1, S)
    p = orc.Problem(dims, grid, (20, 20, 20), 0.01, D, lam)
    f0 = rng.uniform(0, 1, p.shape)
    a = gpu_domain(p, field=f0)
    b = gpu_domain(p, field=f0)
    a.set_tuning(x_sweep=1)
    b.set_tuning(x_sweep=5)
    a.step(3)
    b.step(3)
    assert np.array_equal(a.field(), b.field())


# the 4-byte chunked kernel asked for every S (x_sweep 6), in particular S = 4
# and 8 (whose default chunked kernel moves 16-byte chunks): bucket tails,
# partial last waves (16 rows a wave for S = 4, 8 for S = 8), too few rows
CHUNK4_SHAPES = [
    (3, (200, 13, 3), 4), (3, (184, 9, 2), 8), (3, (161, 5, 4), 4), (3, (8, 17, 3), 8),
    (2, (97, 23, 1), 4), (3, (59, 7, 3), 8), (3, (128, 8, 1), 8), (3, (192, 31, 2), 4),
    (3, (200, 33, 2), 8), (3, (37, 21, 3), 4), (3, (193, 6, 3), 4), (3, (40, 3, 2), 4),
    (3, (200, 9, 5), 3), (3, (100, 7, 9), 2), (3, (120, 4, 7), 5),
]


@pytest.mark.parametrize("dims,grid,S", CHUNK4_SHAPES)
def test_chunked4_x_sweep_matches_strided(dims, grid, S):
    """The 4-byte chunked kernel for any S: the strided kernel's fields bit for
    bit over several steps"""
    rng = np.random.default_rng(hash((dims, grid, S, 6)) % 2**32)
    D = 10 ** rng.uniform(0, 5, S)
    lam = rng.uniform(0.01, 1, S)
    p = orc.Problem(dims, grid, (20, 20, 20), 0.01, D, lam)
    f0 = rng.uniform(0, 1, p.shape)
    a = gpu_domain(p, field=f0)
    b = gpu_domain(p, field=f0)
    a.set_tuning(x_sweep=1)
    b.set_tuning(x_sweep=6)
    a.step(3)
    b.step(3)
    assert np.array_equal(a.field(), b.field())


# workgroups of one, two or four waves for the one-lane register kernels (y/z
# column lines, strided x lines, both chunked x kernels)
LPB_SHAPES = [(3, (200, 13, 3), 4), (3, (61, 37, 9), 3), (2, (97, 23, 1), 8), (3, (33, 40, 5), 1),
              (3, (184, 9, 6), 2), (3, (20, 30, 70), 5)]


@pytest.mark.parametrize("x_sweep", [0, 1, 5, 6])
@pytest.mark.parametrize("dims,grid,S", LPB_SHAPES)
def test_lines_per_block_same_results(dims, grid, S, x_sweep):
    """lines_per_block 64 and 128 change only the workgroup size: the fields
    equal the default's (256) bit for bit over several steps"""
    rng = np.random.default_rng(hash((dims, grid, S, 64)) % 2**32)
    D = 10 ** rng.uniform(0, 5, S)
    lam = rng.uniform(0.01, 1, S)
    p = orc.Problem(dims, grid, (20, 20, 20), 0.01, D, lam)
    f0 = rng.uniform(0, 1, p.shape)
    out = []
    for lpb in (0, 64, 128):
        d = gpu_domain(p, field=f0)
        d.set_tuning(x_sweep=x_sweep, lines_per_block=lpb)
        d.step(3)
        out.append(d.field())
    assert np.array_equal(out[0], out[1]) and np.array_equal(out[0], out[2])


def test_lines_per_block_in_a_batch():
    """A batch whose domains use different workgroup sizes equals each domain
    stepped alone bit for bit."""
    from physicore_amd import bfvm
    rng = np.random.default_rng(64)
    ps = [orc.Problem(3, g, (20, 20, 20), 0.01, list(10 ** rng.uniform(0, 5, S)), list(rng.uniform(0.01, 1, S)))
          for g, S in (((40, 30, 20), 4), ((64, 17, 33), 3), ((48, 48, 16), 4), ((100, 12, 9), 2))]
    f0 = [rng.uniform(0, 1, p.shape) for p in ps]
    alone = [gpu_domain(p, field=f) for p, f in zip(ps, f0)]
    mixed = [gpu_domain(p, field=f) for p, f in zip(ps, f0)]
    for k, (a, m) in enumerate(zip(alone, mixed)):
        lpb = (0, 64, 128, 64)[k]
        a.set_tuning(lines_per_block=lpb)
        m.set_tuning(lines_per_block=lpb)
        a.step(4)
    b = bfvm.Batch(mixed)
    b.step(4)
    b.sync()
    for a, m in zip(alone, mixed):
        assert np.array_equal(a.field(), m.field())


# the x sweep in two halves on two streams (split_xy 1) against one launch per
# axis (split_xy 2); each shape splits both sweeps inside the domain: x kernels
# chunked 4-byte (S 4, rows of 200), chunked 16-byte (S 8), two lanes per line
# (rows of 260), strided lines and the scan; y column lines and two lanes per
# line (y lines of 300)
SPLIT_SHAPES = [((200, 13, 30), 4, 0), ((120, 20, 24), 8, 0), ((260, 16, 20), 4, 0), ((40, 300, 12), 2, 0),
                ((200, 13, 30), 4, 1), ((300, 10, 12), 1, 3), ((61, 37, 9), 3, 0)]


@pytest.mark.parametrize("grid,S,x_sweep", SPLIT_SHAPES)
def test_split_xy_same_results(grid, S, x_sweep):
    """The halves change only the launches: the fields equal one launch per
    axis bit for bit over eager steps and replayed graphs"""
    rng = np.random.default_rng(hash((grid, S, 65)) % 2**32)
    p = orc.Problem(3, grid, (20, 20, 20), 0.01, 10 ** rng.uniform(0, 5, S), rng.uniform(0.01, 1, S))
    f0 = rng.uniform(0, 1, p.shape)
    out = []
    for split in (1, 2):
        d = gpu_domain(p, field=f0)
        d.set_tuning(x_sweep=x_sweep, split_xy=split)
        plan = d.sweep_plan()
        if split == 1:
            assert 0 < plan["split_x"] < plan["x_tiles"] and 0 < plan["split_y"] < plan["y_tiles"], plan
        else:
            assert plan["split_x"] == plan["split_y"] == 0
        d.step(3)
        d.step_graph(20, 10)
        out.append(d.field())
    assert np.array_equal(out[0], out[1])


# the default: halves when the x lanes (the scan kernel excepted) and the y
# lanes each fill more than one round of the register file and the y lines run
# two waves per SIMD
DEFAULT_SPLIT = [((200, 200, 200), 4, True), ((176, 176, 176), 4, False), ((200, 245, 200), 4, True),
                 ((400, 400, 100), 2, True), ((250, 250, 250), 4, True), ((160, 160, 160), 8, True),
                 ((128, 128, 128), 8, False), ((300, 300, 300), 1, True), ((400, 300, 100), 2, False),
                 ((300, 120, 120), 1, False)]


@pytest.mark.parametrize("grid,S,split", DEFAULT_SPLIT)
def test_split_xy_default_choice(grid, S, split):
    from physicore_amd import bfvm
    d = bfvm.Domain(3, grid, (20, 20, 20), 0.01, np.full(S, 1e4), np.full(S, 0.1), initial=np.ones(S))
    plan = d.sweep_plan()
    d.close()
    assert (plan["split_x"] > 0) == split, plan


def test_split_xy_full_step_workload():
    """A whole step (Dirichlet face, bulk sphere, cells) with the halves equals
    one launch per axis bit for bit, and the default (halves where they pay)
    equals both"""
    from physicore_amd import workloads as wl
    w = wl.c2(side=64, n_cells=5000)
    out = []
    for split in (1, 2, 0):
        d = wl.make_domain(w)
        d.set_tuning(split_xy=split)
        d.step(2)
        d.step_graph(10, 10)
        out.append(d.field())
        d.close()
    assert np.array_equal(out[0], out[1]) and np.array_equal(out[0], out[2])


def test_split_xy_rejects_other_values():
    from physicore_amd import bfvm
    d = gpu_domain(orc.Problem(3, (8, 8, 8), (20, 20, 20), 0.01, [1e3], [0.1]))
    with pytest.raises(bfvm.BfvmError):
        d.set_tuning(split_xy=3)


# rows longer than one lane holds (past 176/192/200 voxels for S = 1/5,7/others):
# the two-lane chunked kernel, its bucket edges, partial wave pairs, and the
# longest rows of each S
LONG_X_SHAPES = [
    (3, (400, 40, 1), 2), (3, (201, 70, 2), 1), (3, (384, 12, 3), 3), (3, (230, 13, 2), 5),
    (3, (390, 11, 1), 6), (3, (193, 9, 2), 7), (2, (300, 64, 1), 2), (3, (399, 33, 1), 2),
    (3, (352, 65, 1), 1), (3, (384, 9, 1), 7), (3, (209, 21, 1), 3), (3, (385, 10, 2), 6),
]
# S = 4 and 8 (4-byte chunks, x_sweep 6; S = 8 rows past 256 voxels take it by default)
LONG_X_SHAPES_48 = [
    (3, (250, 20, 3), 4), (3, (400, 16, 2), 4), (3, (201, 9, 2), 8), (3, (384, 8, 2), 8),
    (2, (300, 40, 1), 4), (3, (333, 17, 1), 8), (3, (257, 12, 2), 8), (3, (216, 7, 5), 4),
]


@pytest.mark.parametrize("dims,grid,S", LONG_X_SHAPES + LONG_X_SHAPES_48)
def test_two_lane_x_sweep_matches_oracle(dims, grid, S):
    rng = np.random.default_rng(hash((dims, grid, S, 7)) % 2**32)
    D = 10 ** rng.uniform(0, 5, S)
    lam = rng.uniform(0.01, 1, S)
    p = orc.Problem(dims, grid, (20, 20, 20), 0.01, D, lam)
    f0 = rng.uniform(0, 1, p.shape)
    want = p.diffuse(f0.astype(np.float32).astype(np.float64))
    d = gpu_domain(p, field=f0)
    d.set_tuning(x_sweep=6 if S in (4, 8) else 5)
    d.step(1)
    err = rel_linf(d.field(), want)
    assert err.max() < 2e-6, err


# rows a lane could hold, split over two lanes on request (x_sweep 7): the
# shortest half-row bucket (48), bucket edges, partial wave pairs, every S
SHORT_TWO_LANE_SHAPES = [
    (3, (65, 40, 3), 1), (3, (80, 33, 2), 2), (3, (97, 21, 2), 3), (3, (128, 17, 3), 4),
    (3, (160, 9, 2), 5), (3, (176, 7, 3), 6), (3, (192, 9, 2), 7), (3, (200, 11, 2), 8),
    (2, (96, 70, 1), 4), (3, (113, 13, 3), 8), (3, (66, 19, 5), 5), (3, (200, 16, 2), 2),
]


@pytest.mark.parametrize("dims,grid,S", SHORT_TWO_LANE_SHAPES)
def test_short_rows_on_two_lanes_match_oracle(dims, grid, S):
    rng = np.random.default_rng(hash((dims, grid, S, 17)) % 2**32)
    D = 10 ** rng.uniform(0, 5, S)
    lam = rng.uniform(0.01, 1, S)
    p = orc.Problem(dims, grid, (20, 20, 20), 0.01, D, lam)
    f0 = rng.uniform(0, 1, p.shape)
    want = p.diffuse(f0.astype(np.float32).astype(np.float64))
    d = gpu_domain(p, field=f0)
    d.set_tuning(x_sweep=7)
    d.step(1)
    err = rel_linf(d.field(), want)
    assert err.max() < 2e-6, err


@pytest.mark.parametrize("S", [4, 8])
def test_two_lane_x_sweep_s48_batched_equals_alone(S):
    from physicore_amd import bfvm
    rng = np.random.default_rng(9 + S)
    p = orc.Problem(3, (310, 24, 3), (20, 20, 20), 0.01, list(10 ** rng.uniform(0, 5, S)),
                    list(rng.uniform(0.01, 1, S)))
    f0 = rng.uniform(0, 1, p.shape)
    a, b = gpu_domain(p, field=f0), gpu_domain(p, field=f0)
    for d in (a, b):
        d.set_tuning(x_sweep=6)
    batch = bfvm.Batch([b])
    a.step(4)
    batch.step(4)
    assert np.array_equal(a.field(), b.field())


def test_two_lane_x_sweep_batched_equals_alone():
    from physicore_amd import bfvm
    rng = np.random.default_rng(8)
    p = orc.Problem(3, (400, 40, 3), (20, 20, 20), 0.01, [1e4, 30.0], [0.1, 0.5])
    f0 = rng.uniform(0, 1, p.shape)
    a, b = gpu_domain(p, field=f0), gpu_domain(p, field=f0)
    for d in (a, b):
        d.set_tuning(x_sweep=5)
    batch = bfvm.Batch([b])
    a.step(4)
    batch.step(4)
    assert np.array_equal(a.field(), b.field())


def test_fp32_arithmetic_mode_runs():
    rng = np.random.default_rng(4)
    p = orc.Problem(3, (20, 30, 40), (20, 20, 20), 0.01, [1e3, 3.0], [0.1, 0.5])
    f0 = rng.uniform(0, 1, p.shape)
    want = p.diffuse(f0.astype(np.float32).astype(np.float64))
    d = gpu_domain(p, field=f0, precision="fp32")
    d.step(1)
    assert rel_linf(d.field(), want).max() < 1e-5


def _fp32_neighbours(x):
    """The FP32 values just below and above each FP64 value (x not representable)."""
    lo = x.astype(np.float32)
    lo = np.where(lo.astype(np.float64) > x, np.nextafter(lo, np.float32(-np.inf)), lo)
    return lo, np.nextafter(lo, np.float32(np.inf))


def test_stochastic_rounding_distribution():
    """With D = 0 a 1-D step is one FP64 product c0*b per voxel, rounded once
    to FP32 (lines of 1000: the LDS-tile sweep).  Stochastic rounding puts
    each value on one of its two FP32 neighbours and rounds up with
    probability equal to the discarded fraction of the ulp; round_nearest
    gives numpy's FP32 rounding."""
    lam = np.array([0.7, 0.3, 0.9, 0.5])
    p = orc.Problem(1, (1000, 1, 1), (20, 20, 20), 0.01, np.zeros(4), lam)
    c0 = np.float32(0.3)
    field = np.full(p.shape, float(c0))
    exact = float(c0) / (1 + p.dt * lam)  # b of a decay-only line, dims = 1
    lo, hi = _fp32_neighbours(exact)
    frac = (exact - lo.astype(np.float64)) / (hi.astype(np.float64) - lo.astype(np.float64))
    d = gpu_domain(p, field=field)
    d.step(1)
    f = d.field().reshape(-1, 4)
    for s in range(4):
        assert np.all((f[:, s] == lo[s]) | (f[:, s] == hi[s])), s
        up = np.mean(f[:, s] == hi[s])
        assert abs(up - frac[s]) < 5 * np.sqrt(frac[s] * (1 - frac[s]) / len(f)) + 1e-3, (s, up, frac[s])
    d = gpu_domain(p, field=field)
    d.set_tuning(round_nearest=1)
    d.step(1)
    assert np.array_equal(d.field().reshape(-1, 4), np.broadcast_to(exact.astype(np.float32), (1000, 4)))


@pytest.mark.parametrize("last", ["none", "bulk", "cells", "internalized"])
def test_step_count_advances(last):
    """The rounding keys change every step whichever stage ends it: a domain
    restarted from another's field after one step, with its step count at 0,
    rounds its next step differently; the uninterrupted run does not."""
    internalized = last == "internalized"
    if internalized:
        p, f0, dirs, bulk, ag = sc.config4(seed=12, n_cells=4000, nx=24, ny=24, nz=12)
    else:
        p, f0, dirs, bulk, ag = sc.config2(side=24, S=3, n_cells=400, seed=9, with_cells=last == "cells",
                                           with_bulk=last == "bulk")

    def make(field):
        return gpu_domain(p, field=field, dirichlet=dirs, bulk=bulk, agents=ag, internalized=internalized,
                          reduce="sorted")

    a, c = make(f0), make(f0)
    a.step(1)
    b = make(a.field())
    a.step(1)
    b.step(1)
    c.step(2)
    assert np.array_equal(a.field(), c.field())
    assert not np.array_equal(a.field(), b.field())


def test_stochastic_rounding_unbiased_3d():
    """Three roundings per step (x scan, y and z register lines): E[field] is
    the exact c0*b^3, so the voxel mean matches it to a few standard errors
    (round-to-nearest is off by a fixed fraction of an ulp), and every voxel is
    within 3 ulps of it."""
    lam = 0.7
    p = orc.Problem(3, (64, 64, 64), (20, 20, 20), 0.01, [0.0], [lam])
    c0 = np.float32(0.3)
    exact = float(c0) / (1 + p.dt * lam / 3) ** 3
    ulp = float(np.spacing(np.float32(exact)))
    d = gpu_domain(p, field=np.full(p.shape, float(c0)))
    d.step(1)
    f = d.field().astype(np.float64).ravel()
    assert np.abs(f - exact).max() <= 3 * ulp
    assert len(np.unique(f)) > 1
    sigma = 0.5 * np.sqrt(3) * ulp / np.sqrt(f.size)
    assert abs(f.mean() - exact) < 6 * sigma, (f.mean() - exact) / ulp


def test_pure_diffusion_mass_balance():
    """Neumann sweeps conserve line mass up to the decay share, so k steps divide
    each substrate's total by (1 + dt*lambda/3)^(3k) exactly."""
    rng = np.random.default_rng(6)
    p = orc.Problem(3, (48, 40, 56), (20, 20, 20), 0.01, [1e5, 10.0, 1.0], [0.5, 0.01, 1.0])
    d = gpu_domain(p, field=rng.uniform(0, 1, p.shape))
    m0 = d.total_mass()
    k = 50
    d.step(k)
    m1 = d.total_mass()
    np.testing.assert_allclose(m1, m0 / (1 + p.dt * p.decay / 3) ** (3 * k), rtol=2e-6)
    # the device reduction equals an FP64 host sum of the same FP32 field
    f = d.field(np.float64)
    np.testing.assert_allclose(m1, f.sum(axis=(0, 1, 2)) * 8000.0, rtol=1e-12)


# ----------------------------------------------------------- Dirichlet / bulk

@pytest.mark.parametrize("name,dims", [("Interior1D", 1), ("Interior2D", 2), ("Interior3D", 3)])
def test_dirichlet_interior(name, dims):
    t = G["dirichlet"][name]
    env = G["dirichlet"]["default_microenv"]
    p = sc.problem_from_env(env, dims, t["bbox_max"], t["voxel"])
    # D = lambda = 0 makes the sweeps the identity, so a step is the Dirichlet stage
    q = orc.Problem(dims, (p.nx, p.ny, p.nz), p.voxel, p.dt, [0.0, 0.0], [0.0, 0.0])
    spec = orc.DirichletSpec(2, dims)
    spec.set_interior([v[:dims] for v in t["voxels"]], [[v, 0] for v in t["values"]],
                      [[1, 0] for _ in t["values"]])
    d = gpu_domain(q, field=sc.constant_field(q, env["initial"]), dirichlet=spec)
    d.step(1)
    f = d.field(np.float64)
    for v, val in zip(t["voxels"], t["values"]):
        x, y, z = v
        assert f[z, y, x, 0] == val
    assert np.all(f[..., 1] == 1)


@pytest.mark.parametrize("dims", [1, 2, 3])
def test_dirichlet_boundary_precedence(dims):
    t = G["dirichlet"]["Boundary"]
    env = G["dirichlet"]["default_microenv"]
    p = sc.problem_from_env(env, dims, t["bbox_max"], t["voxel"])
    q = orc.Problem(dims, (p.nx, p.ny, p.nz), p.voxel, p.dt, [0.0, 0.0], [0.0, 0.0])
    spec = orc.DirichletSpec(2, dims)
    for a, ax in enumerate("xyz"[:dims]):
        lo, hi = t["face_values"][ax]
        spec.set_face(a, "min", [lo, 42], [1, 0])
        spec.set_face(a, "max", [hi, 42], [1, 0])
    want = orc.dirichlet(q, sc.constant_field(q, env["initial"]), spec)
    d = gpu_domain(q, field=sc.constant_field(q, env["initial"]), dirichlet=spec)
    d.step(1)
    assert np.array_equal(d.field(np.float64), want)


def test_dirichlet_precedence_interior_over_faces_random():
    rng = np.random.default_rng(8)
    S = 3
    p = orc.Problem(3, (9, 8, 7), (10, 10, 10), 0.05, [50.0, 5.0, 500.0], [0.1, 0.2, 0.3])
    spec = orc.DirichletSpec(S, 3)
    for a in range(3):
        for side in ("min", "max"):
            spec.set_face(a, side, rng.uniform(0, 9, S), rng.integers(0, 2, S))
    vox = [(0, 0, 0), (8, 7, 6), (4, 4, 4), (0, 3, 6), (4, 4, 4)]
    spec.set_interior(vox, rng.uniform(10, 20, (5, S)), rng.integers(0, 2, (5, S)))
    f0 = rng.uniform(0, 1, p.shape)
    s = orc.Solver(p, f0.astype(np.float32).astype(np.float64), dirichlet=spec)
    s.solve(3)
    d = gpu_domain(p, field=f0, dirichlet=spec)
    d.step(3)
    assert rel_linf(d.field(), s.field).max() < 3e-6


@pytest.mark.parametrize("dims,grid", [(3, (6, 5, 301)), (2, (7, 333, 1)), (3, (4, 290, 3)), (3, (6, 5, 240)),
                                       (2, (7, 220, 1)), (3, (5, 231, 4))])
def test_long_lines_with_faces(dims, grid):
    """Lines of 209..512 voxels (two lanes per line) in the sweep that applies
    the Dirichlet faces (z in 3-D, y in 2-D), every face set for some
    substrate, and lines lying in faces of the other axes."""
    rng = np.random.default_rng(sum(grid))
    S = 2
    p = orc.Problem(dims, grid, (10, 10, 10), 0.01, [800.0, 3.0], [0.1, 0.6])
    spec = orc.DirichletSpec(S, dims)
    for a in range(dims):
        for side in ("min", "max"):
            spec.set_face(a, side, rng.uniform(0, 9, S), rng.integers(0, 2, S))
    f0 = rng.uniform(0, 1, p.shape)
    s = orc.Solver(p, f0.astype(np.float32).astype(np.float64), dirichlet=spec)
    s.solve(3)
    d = gpu_domain(p, field=f0, dirichlet=spec)
    d.step(3)
    assert rel_linf(d.field(), s.field).max() < 3e-6


def test_bulk_golden():
    t = G["bulk"]["SimulateBulkSource"]
    p = sc.problem_from_env(t, 3, t["bbox_max"], t["voxel"])
    q = orc.Problem(3, (p.nx, p.ny, p.nz), p.voxel, p.dt, [0.0, 0.0], [0.0, 0.0])
    sup, upt, tgt = (np.zeros(q.shape) for _ in range(3))
    x, y, z = t["point"]
    sup[z, y, x, 0], tgt[z, y, x, 0], upt[z, y, x, 0] = t["supply"], t["target"], t["uptake"]
    d = gpu_domain(q, field=sc.constant_field(q, t["initial"]), bulk=(sup, upt, tgt))
    d.step(1)
    f = d.field(np.float64)
    assert abs(f[z, y, x, 0] - t["expected_point"]) < t["tol_point"]
    f[z, y, x, 0] = 10
    assert np.all(f[..., 0] == 10) and np.all(f[..., 1] == 1)


def test_bulk_dense_random_matches_oracle():
    rng = np.random.default_rng(9)
    p = orc.Problem(3, (12, 10, 8), (20, 20, 20), 0.01, [100.0, 1.0], [0.1, 0.3])
    sup, upt, tgt = (np.where(rng.uniform(0, 1, p.shape) < 0.3, rng.uniform(0, 5, p.shape), 0)
                     for _ in range(3))
    f0 = rng.uniform(0, 1, p.shape)
    s = orc.Solver(p, f0.astype(np.float32).astype(np.float64), bulk=(sup, upt, tgt))
    s.solve(2)
    d = gpu_domain(p, field=f0, bulk=(sup, upt, tgt))
    d.step(2)
    assert rel_linf(d.field(), s.field).max() < 2e-6


@pytest.mark.parametrize("reduce", ["atomic", "sorted"])
def test_bulk_then_cells_closed_form(reduce):
    """A voxel in the bulk list that holds a cell gets cells(bulk(D)), the stage
    order of openmp_solver.cpp:27-37 (test_oracle pins the same closed form on
    the oracle); alone the plain step runs bulk and cells as one launch."""
    import test_oracle as to

    p, f, bulk_arrays, ag, bulk, cell = to._bulk_and_cell_voxel_case()
    d = gpu_domain(p, field=f, bulk=bulk_arrays, agents=ag, reduce=reduce)
    d.step(1)
    got = d.field(np.float64).reshape(-1)
    want = cell(bulk(5.0))
    assert abs(got[1] - want) <= 2e-7 * abs(want), (got[1], want, bulk(cell(5.0)))
    assert got[0] == 5.0 and got[2] == 5.0


# ---------------------------------------------------------------------- cells

def _cells_problem(p):
    return orc.Problem(p.dims, (p.nx, p.ny, p.nz), p.voxel, p.dt, [0.0] * p.S, [0.0] * p.S)


@pytest.mark.parametrize("reduce", ["atomic", "sorted"])
@pytest.mark.parametrize("dims", [1, 2, 3])
@pytest.mark.parametrize("internalized", [True, False])
@pytest.mark.parametrize("recompute", [True, False])
def test_cells_simple_golden(reduce, dims, internalized, recompute):
    g = G["cells"]["Simple"]
    p, ag = sc.simple_cells(dims)
    q = _cells_problem(p)
    d = gpu_domain(q, field=sc.constant_field(q, [1, 1]), agents=ag, internalized=internalized,
                   reduce=reduce)
    diag = [(0, 0, 0), (1, 1, 1), (2, 2, 2)]

    def check(step):
        f = d.field(np.float64)
        I = d.internalized()
        for k in range(3):
            x, y, z = diag[k]
            for sub in range(2):
                v = f[z if dims > 2 else 0, y if dims > 1 else 0, x, sub]
                want, tol = g[f"step{step}_density"][2 * k + sub]
                assert abs(v - want) < max(tol, 3e-7 * abs(want))
                if internalized:
                    want, tol = g[f"step{step}_internalized"][2 * k + sub]
                    assert abs(I[k, sub] - want) < max(tol, 2e-7 * abs(want))

    d.step(1)
    check(1)
    if recompute:
        d.recompute_positional_data()
    d.step(1)
    check(2)
    d.release_internalized(0)
    if internalized:
        # the reference gets exactly 1 in FP64; the FP32 field is within its rounding
        f = d.field(np.float64)
        assert abs(f[0, 0, 0, 0] - 1.0) < 1e-6 and abs(f[0, 0, 0, 1] - 1.0) < 1e-6
        assert np.all(d.internalized()[0] == 0)


@pytest.mark.parametrize("builder", [sc.conflict_cells, sc.conflict_big_cells])
@pytest.mark.parametrize("reduce", ["atomic", "sorted"])
@pytest.mark.parametrize("internalized", [True, False])
def test_cells_conflict_match_oracle(builder, reduce, internalized):
    p, ag = builder()
    q = _cells_problem(p)
    f_or = sc.constant_field(q, [1, 1])
    ag_or = orc.AgentSpec(ag.n, 2, 1)
    for k in ("positions", "secretion", "saturation", "uptake", "net_export", "fraction_released",
              "volumes"):
        setattr(ag_or, k, getattr(ag, k).copy())
    s = orc.Solver(q, f_or, agents=ag_or, compute_internalized=internalized)
    d = gpu_domain(q, field=sc.constant_field(q, [1, 1]), agents=ag, internalized=internalized,
                   reduce=reduce)
    steps, peak = 2, 1.0
    for _ in range(steps):
        s.solve(1)
        d.step(1)
        assert rel_linf(d.field(), f_or).max() < 1e-6
        peak = max(peak, float(np.abs(f_or).max()))
        if internalized:
            # the GPU reads D_old from the FP32 field, the oracle from FP64
            assert rel_linf(d.internalized(), ag_or.internalized).max() < 1e-6
    for i in range(ag.n):
        s.release(i)
        d.release_internalized(i)
    # The FP64 oracle returns to the initial density exactly.  The FP32 field
    # stores each intermediate density once per step, stochastically rounded
    # (error below one ulp, <= 2^-23 * peak; the densities here reach ~1.2e3),
    # and the internalized amounts are FP64 differences of the unrounded
    # values, so the release is off by at most that rounding per step.
    assert np.abs(d.field() - f_or).max() <= 1e-6 + steps * peak * 2.0 ** -23


@pytest.mark.parametrize("reduce", ["atomic", "sorted"])
@pytest.mark.parametrize("shuffled", [False, True])
def test_internalized_lattice_with_one_shared_voxel(reduce, shuffled):
    """A lattice of agents, one per voxel in voxel order, plus two agents
    sharing voxels (so every agent takes the per-agent update, as in the
    reference's benchmark): the owners alone in their voxel are updated by
    the cell apply when the owner list is in agent order, and by the per-agent
    pass when the agents are shuffled; fields and internalized amounts match
    the oracle either way over steps with a recompute between them"""
    rng = np.random.default_rng(21)
    p = orc.Problem(3, (20, 16, 12), (20, 20, 20), 0.01, [1000.0, 50.0, 5.0], [0.1, 0.3, 0.02])
    n_lat = 20 * 16 * 10
    idx = np.arange(n_lat)
    pos = np.stack([idx % 20, (idx // 20) % 16, idx // 320], axis=1) * 20.0 + 10.0
    pos = np.concatenate([pos, pos[[0, 777]] + 1.0])
    n = len(pos)
    if shuffled:
        pos = pos[rng.permutation(n)]
    ag = orc.AgentSpec(n, p.S, 3)
    ag.positions[:] = pos
    ag.secretion[:] = rng.uniform(0, 10, (n, p.S))
    ag.uptake[:] = rng.uniform(0, 10, (n, p.S))
    ag.saturation[:] = rng.uniform(0, 2, (n, p.S))
    ag.net_export[:] = rng.uniform(-1, 1, (n, p.S))
    ag.fraction_released[:] = rng.uniform(0, 1, (n, p.S))
    ag.volumes[:] = rng.uniform(1000, 3000, n)
    f0 = rng.uniform(0, 1, p.shape)
    s = orc.Solver(p, f0.astype(np.float32).astype(np.float64), agents=ag, compute_internalized=True)
    d = gpu_domain(p, field=f0, agents=ag, internalized=True, reduce=reduce)
    for k in range(4):
        if k == 2:
            s.recompute()
            d.recompute_positional_data()
        s.solve(1)
        d.step(1)
        assert rel_linf(d.field(), s.field).max() < 2e-6
        assert rel_linf(d.internalized(), ag.internalized).max() < 2e-6


def _random_agents(p, n, rng, clustered=False):
    ag = orc.AgentSpec(n, p.S, p.dims)
    box = np.array([p.nx * p.voxel[0], p.ny * p.voxel[1], p.nz * p.voxel[2]], dtype=np.float64)
    if clustered:
        pos = rng.normal(box / 2, box / 12, (n, 3))
        pos = np.clip(pos, 0, np.nextafter(box, 0))
    else:
        pos = rng.uniform(0, box, (n, 3))
    ag.positions[:] = pos[:, : p.dims]
    ag.secretion[:] = rng.uniform(0, 10, (n, p.S))
    ag.uptake[:] = rng.uniform(0, 10, (n, p.S))
    ag.saturation[:] = rng.uniform(0, 2, (n, p.S))
    ag.net_export[:] = rng.uniform(-1, 1, (n, p.S))
    ag.fraction_released[:] = rng.uniform(0, 1, (n, p.S))
    ag.volumes[:] = rng.uniform(1000, 3000, n)
    ag.is_active[:] = rng.uniform(0, 1, n) < 0.95
    return ag


@pytest.mark.parametrize("reduce", ["atomic", "sorted"])
@pytest.mark.parametrize("clustered", [False, True])
def test_cell_sums_match_oracle(reduce, clustered):
    rng = np.random.default_rng(11)
    p = orc.Problem(3, (30, 20, 10), (10, 10, 10), 0.01, [0.0, 0.0, 0.0], [0.0, 0.0, 0.0])
    ag = _random_agents(p, 20000, rng, clustered)
    s = orc.Solver(p, np.ones(p.shape), agents=ag)
    s.solve(1)
    num, den, fac, ballots, conflict = s.reduced()
    d = gpu_domain(p, field=np.ones(p.shape), agents=ag, reduce=reduce)
    owners, gn, gd, gf, gconf = d.cell_sums()
    want_owners = np.unique(ballots[ballots >= 0])
    if reduce == "sorted":
        # same owners (lowest active index per voxel) and the same FP64 sums, bit for bit
        assert np.array_equal(owners, want_owners)
        assert np.array_equal(gn, num[owners]) and np.array_equal(gd, den[owners])
        assert np.array_equal(gf, fac[owners])
    else:
        # atomic election may pick another owner; compare per voxel
        vox_of = lambda o: (ag.positions[o] // 10).astype(int) @ np.array([1, 30, 600])
        gmap = {int(v): k for k, v in enumerate(vox_of(owners))}
        assert len(gmap) == len(want_owners)
        for k, o in enumerate(want_owners):
            j = gmap[int(vox_of(np.array([o]))[0])]
            np.testing.assert_allclose(gn[j], num[o], rtol=1e-13, atol=1e-300)
            np.testing.assert_allclose(gd[j], den[o], rtol=1e-13)
            np.testing.assert_allclose(gf[j], fac[o], rtol=1e-12, atol=1e-18)
    assert gconf == conflict


def test_sorted_mode_is_reproducible():
    rng = np.random.default_rng(12)
    p, f0, dirs, bulk, _ = sc.config2(side=40, S=3, n_cells=0, seed=5)
    ag = _random_agents(p, 30000, rng, clustered=True)
    fields = []
    for _ in range(2):
        d = gpu_domain(p, field=f0, dirichlet=dirs, bulk=bulk, agents=ag, internalized=True,
                       reduce="sorted")
        d.step(5)
        fields.append((d.field(), d.internalized()))
    assert np.array_equal(fields[0][0], fields[1][0])
    assert np.array_equal(fields[0][1], fields[1][1])


def test_cells_inactive_and_reactivation():
    p, ag = sc.simple_cells(1)
    q = _cells_problem(p)
    ag.is_active[1] = 0
    d = gpu_domain(q, field=sc.constant_field(q, [1, 1]), agents=ag, internalized=True)
    d.step(1)
    f = d.field(np.float64)
    assert abs(f[0, 0, 0, 0] - 28) < 1e-5 and f[0, 0, 1, 0] == 1 and f[0, 0, 1, 1] == 1
    assert np.all(d.internalized()[1] == 0)
    # all inactive: nothing changes
    ag.is_active[:] = 0
    d2 = gpu_domain(q, field=sc.constant_field(q, [1, 1]), agents=ag, internalized=True)
    d2.step(1)
    assert np.all(d2.field() == 1) and np.all(d2.internalized() == 0)


def test_agent_outside_grid_is_an_error():
    """The recompute runs on the stream without a host wait: the error comes
    from the next call that waits for it, and the positions stay marked for
    another recompute."""
    p, ag = sc.simple_cells(3)
    ag.positions[2] = [60.0, 10.0, 10.0]  # on the max face: index n
    d = gpu_domain(_cells_problem(p), field=np.ones(p.shape), agents=ag)
    d.step(1)
    with pytest.raises(Exception) as e:
        d.sync()
    assert "outside" in str(e.value)
    d.step(1)
    with pytest.raises(Exception) as e:
        d.field()
    assert "outside" in str(e.value)


def _hip_copy_to_device(dst_ptr, arr):
    """Blocking host-to-device copy through the HIP runtime (the domain's stream
    must be idle: call sync first)."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    a = np.ascontiguousarray(arr, dtype=np.float64)
    assert hip.hipMemcpy(dst_ptr, a.ctypes.data, a.nbytes, 1) == 0  # hipMemcpyHostToDevice


@pytest.mark.parametrize("reduce", ["atomic", "sorted"])
def test_positions_moved_on_device_equal_new_agent_arrays(reduce):
    """Agents moved by writing the device positions and recomputing give the
    field of uploading the moved agents, with graphs captured before the move
    kept (the recompute changes no launch)."""
    rng = np.random.default_rng(31)
    p, f0, dirs, bulk, _ = sc.config2(side=36, S=3, n_cells=0, seed=31)
    ag = _random_agents(p, 4000, rng, clustered=True)
    moved = np.clip(ag.positions + rng.normal(0, 30, ag.positions.shape), 0,
                    np.nextafter(np.array([p.nx, p.ny, p.nz]) * 20.0, 0))
    a = gpu_domain(p, field=f0, dirichlet=dirs, bulk=bulk, agents=ag, reduce=reduce)
    b = gpu_domain(p, field=f0, dirichlet=dirs, bulk=bulk, agents=ag, reduce=reduce)
    a.step_graph(10, 5)
    b.step_graph(10, 5)
    ag.positions[:] = moved
    set_agents(a, ag)
    b.sync()
    _hip_copy_to_device(b.agent_positions_device_ptr(), moved)
    b.recompute_positional_data()
    a.step_graph(10, 5)
    b.step_graph(10, 5)
    fa, fb = a.field(), b.field()
    if reduce == "sorted":
        assert np.array_equal(fa, fb)
    else:
        assert rel_linf(fb, fa.astype(np.float64)).max() < 1e-6
    sa, sb = a.cell_sums(), b.cell_sums()
    assert np.array_equal(sa[0], sb[0]) or reduce == "atomic"
    assert len(sa[0]) == len(sb[0]) and sa[4] == sb[4]


def test_positions_moved_on_device_outside_grid_is_reported():
    p, ag = sc.simple_cells(3)
    d = gpu_domain(_cells_problem(p), field=np.ones(p.shape), agents=ag)
    d.step(1)
    d.sync()
    bad = ag.positions.copy()
    bad[1] = [-5.0, 10.0, 10.0]
    _hip_copy_to_device(d.agent_positions_device_ptr(), bad)
    d.recompute_positional_data()
    d.step(1)
    with pytest.raises(Exception) as e:
        d.sync()
    assert "outside" in str(e.value)


# --------------------------------------------------------------- full steps

def _full_run(p, f0, dirs, bulk, ag, steps, internalized, reduce="atomic", graph=0):
    ag_or = None
    if ag is not None:
        ag_or = orc.AgentSpec(ag.n, p.S, p.dims)
        for k in ("positions", "secretion", "saturation", "uptake", "net_export",
                  "fraction_released", "volumes", "is_active"):
            setattr(ag_or, k, getattr(ag, k).copy())
    f_or = f0.astype(np.float32).astype(np.float64)
    s = orc.Solver(p, f_or, dirichlet=dirs, bulk=bulk, agents=ag_or,
                   compute_internalized=internalized)
    s.solve(steps)
    d = gpu_domain(p, field=f0, dirichlet=dirs, bulk=bulk, agents=ag, internalized=internalized,
                   reduce=reduce)
    if graph:
        d.step_graph(steps, graph)
    else:
        d.step(steps)
    return d, f_or, ag_or


def test_config1_full_100_steps():
    p, f0, dirs, bulk, ag = sc.config1()
    d, f_or, _ = _full_run(p, f0, dirs, bulk, ag, 100, internalized=False)
    assert rel_linf(d.field(), f_or).max() < 1e-5


@pytest.mark.parametrize("reduce", ["atomic", "sorted"])
def test_config2_shape_reduced_grid(reduce):
    """Config-2 physics (x-min Dirichlet, bulk sphere, uniform cells, 4 substrates
    with D in [1, 1e5]) on a 40^3 grid for 60 steps, with internalized tracking."""
    p, f0, dirs, bulk, ag = sc.config2(side=40, S=4, n_cells=4000, seed=7)
    d, f_or, ag_or = _full_run(p, f0, dirs, bulk, ag, 60, internalized=True, reduce=reduce)
    assert rel_linf(d.field(), f_or).max() < 1e-5
    assert rel_linf(d.internalized(), ag_or.internalized).max() < 1e-5


def test_graph_replay_matches_eager():
    """A 10-step hipGraph replayed several times, with eager work and syncs in
    between, equals the same steps launched eagerly."""
    p, f0, dirs, bulk, ag = sc.config2(side=32, S=2, n_cells=2000, seed=9)
    # the sorted reduction makes both domains' cell sums bit-identical
    a = gpu_domain(p, field=f0, dirichlet=dirs, bulk=bulk, agents=ag, internalized=True, reduce="sorted")
    b = gpu_domain(p, field=f0, dirichlet=dirs, bulk=bulk, agents=ag, internalized=True, reduce="sorted")
    for _ in range(3):
        a.step(10)
        b.step_graph(10, 10)
        b.sync()
        assert np.array_equal(a.field(), b.field())
        assert np.array_equal(a.internalized(), b.internalized())
        a.total_mass()
        b.total_mass()
    a.step(7)
    b.step_graph(7, 10)  # remainder steps run eagerly
    assert np.array_equal(a.field(), b.field())


def test_persistent_grid_same_result():
    p, f0, dirs, bulk, ag = sc.config2(side=36, S=3, n_cells=1000, seed=10)
    a = gpu_domain(p, field=f0, dirichlet=dirs, bulk=bulk, agents=ag)
    b = gpu_domain(p, field=f0, dirichlet=dirs, bulk=bulk, agents=ag)
    b.set_tuning(persistent_wgs_per_cu=2)
    a.step(4)
    b.step(4)
    assert np.array_equal(a.field(), b.field())


def test_config2_full_size_short():
    """The 200^3 x 4 domain of the benchmark, 3 steps against the FP64 oracle."""
    p, f0, dirs, bulk, ag = sc.config2(side=200, S=4, n_cells=100_000, seed=43)
    d, f_or, _ = _full_run(p, f0, dirs, bulk, ag, 3, internalized=False)
    assert rel_linf(d.field(), f_or).max() < 2e-6


def test_time_sweep_reports_each_axis():
    """bfvm_time_sweep: a positive average for every axis of the domain; an
    axis beyond its dims is rejected."""
    from physicore_amd import bfvm

    require_gpu()
    p = orc.Problem(3, (32, 24, 16), (20, 20, 20), 0.01, [100.0, 10.0], [0.1, 0.2])
    d = gpu_domain(p, field=np.ones(p.shape))
    for axis in range(3):
        assert d.time_sweep(axis, 5) > 0
    with pytest.raises(bfvm.BfvmError):
        d.time_sweep(3, 1)
    q = orc.Problem(2, (32, 24, 1), (20, 20, 20), 0.01, [100.0], [0.1])
    e = gpu_domain(q, field=np.ones(q.shape))
    with pytest.raises(bfvm.BfvmError):
        e.time_sweep(2, 1)


# ------------------------------------------------------------------ reused device blocks

@pytest.mark.parametrize("reduce", ["atomic", "sorted"])
def test_domain_on_reused_blocks_equals_first_run(reduce):
    """A destroyed domain's device blocks are kept for later allocations.  A
    domain built on the blocks of one that held very different values (field
    near 1e6, large rates, many cells) steps to the same field as the first
    build of it: nothing read before it is written."""
    from physicore_amd import serve

    rng = np.random.default_rng(77)
    req = serve.poisson_stream(rng, 1, 1.0, side_range=(40, 40), s_range=(3, 3), max_cells=3000,
                               step_range=(5, 5))[0]
    req.reduce = reduce
    if req.agents is None:
        pytest.skip("stream drew no cells")
    bulk_vox = np.arange(0, req.voxels, 7, dtype=np.int64)
    req.bulk = (bulk_vox, rng.uniform(0, 1, (len(bulk_vox), req.S)), rng.uniform(0, 1, (len(bulk_vox), req.S)),
                rng.uniform(0, 1, (len(bulk_vox), req.S)))

    def run(r):
        d = serve.domain_for(r)
        d.step(r.steps)
        f = d.field(np.float64)
        d.close()
        return f

    first = run(req)
    poison = serve.Request(**{**req.__dict__, "times": {}})
    poison.field = rng.uniform(5e5, 1e6, req.field.shape)
    poison.agents = {k: (None if v is None else np.array(v, copy=True)) for k, v in req.agents.items()}
    poison.agents["secretion"] *= 1e4
    poison.agents["uptake"] *= 1e4
    for _ in range(2):
        run(poison)
        again = run(req)
        if reduce == "sorted":
            assert np.array_equal(again, first)
        else:
            assert rel_linf(again, first).max() < 1e-6


# ------------------------------------------------------------------ uniform bulk in the last sweep

def _bulk_voxels(grid, dims, shape):
    """Listed voxels: every voxel ("dense"), an ellipsoid filling 0.9 of the
    mesh along each axis ("ball": one run per line of the last axis, over an
    eighth of the voxels), that ellipsoid with its middle plane across the last
    axis removed ("split": two runs on some lines), or one filling 0.3 of each
    axis ("small": under an eighth of the voxels)."""
    nx, ny, nz = grid
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    lin = ((z * ny + y) * nx + x).ravel()
    if shape == "dense":
        return lin
    c = [(g - 1) / 2 for g in grid]
    r = [max((0.15 if shape == "small" else 0.45) * g, 0.5) for g in grid]
    q = ((x - c[0]) / r[0]) ** 2 + ((y - c[1]) / r[1]) ** 2 + (((z - c[2]) / r[2]) ** 2 if dims == 3 else 0)
    keep = q.ravel() <= 1.0
    if shape == "split":
        last = (z if dims == 3 else y).ravel()
        keep &= last != int(round(c[dims - 1]))
    return lin[keep]


def _uniform_bulk_workloads(dims, grid, S, seed, shape="dense"):
    """Two workloads alike but for the bulk: the first lists its voxels with
    the same rates (the last sweep applies them when each line holds one run),
    the second the same list with another supply rate at one listed voxel j
    (the bulk stage applies it).  Faces, cells and field are shared.  Returns
    (first, second, j)."""
    from physicore_amd import workloads as wl

    rng = np.random.default_rng(seed)
    V = int(np.prod(grid))
    vox = _bulk_voxels(grid, dims, shape)
    D = 10 ** rng.uniform(0, 4, S)
    lam = rng.uniform(0.01, 1, S)
    field = rng.uniform(0, 1, (grid[2], grid[1], grid[0], S))
    vals = rng.uniform(0, 1, (6, S))
    conds = (rng.uniform(0, 1, (6, S)) < 0.5).astype(np.uint8)
    conds[2 * dims:] = 0
    rates = [np.broadcast_to(rng.uniform(lo, hi, S), (len(vox), S)).copy() for lo, hi in ((0, 1), (0, 1), (0, 2))]
    n = 40
    box = np.asarray(grid[:dims], dtype=np.float64) * 10.0
    agents = dict(positions=rng.uniform(0, 1, (n, dims)) * box, secretion=rng.uniform(0, 5, (n, S)),
                  saturation=np.ones((n, S)), uptake=rng.uniform(0, 5, (n, S)), net_export=np.zeros((n, S)),
                  internalized=None, fraction_released=np.zeros((n, S)), volumes=np.full(n, 500.0),
                  is_active=np.ones(n, dtype=np.uint8))
    k = int(rng.integers(0, len(vox)))
    out = []
    for uniform in (True, False):
        sup = rates[0].copy()
        if not uniform:
            sup[k] *= 1.5
        out.append(wl.Workload(f"uniform bulk {grid}x{S}", dims, tuple(grid), (10, 10, 10), 0.01, D, lam, 5,
                               field=field, faces=(vals, conds),
                               bulk=(vox.astype(np.int64), sup, rates[1], rates[2]), agents=agents))
    assert V >= len(vox)
    return out[0], out[1], int(vox[k])


@pytest.mark.parametrize("nearest", [0, 1])
@pytest.mark.parametrize("shape", ["dense", "ball", "split", "small"])
@pytest.mark.parametrize("dims,grid,S", [(3, (9, 7, 12), 3), (2, (13, 11, 1), 2), (3, (24, 20, 40), 4),
                                         (3, (6, 5, 250), 1)])
def test_uniform_bulk_applied_by_last_sweep(dims, grid, S, shape, nearest):
    """A bulk list with the same rates at every listed voxel, one run of
    listed voxels per line of the last axis and at least an eighth of the
    voxels listed (every voxel, a large ball) is applied by the last sweep as
    it stores its results: no bulk launch follows (a batch of that domain has
    one stage launch fewer), and one step equals the bulk stage's arithmetic
    bit for bit at every voxel but the one whose rates differ in the
    comparison domain.  With two runs on a line, or a small ball, the bulk
    stage runs."""
    from physicore_amd import bfvm
    from physicore_amd import workloads as wl

    wu, wo, j = _uniform_bulk_workloads(dims, grid, S, seed=sum(grid) * 7 + S, shape=shape)
    a, b = wl.make_domain(wu, reduce="sorted"), wl.make_domain(wo, reduce="sorted")
    for d in (a, b):
        d.set_tuning(round_nearest=nearest)
    ba, bb = bfvm.Batch([a]), bfvm.Batch([b])
    ia, ib = ba.info(), bb.info()
    ba.close()
    bb.close()
    fused = shape in ("dense", "ball")
    assert ia["stage_launches"] == ib["stage_launches"] - (1 if fused else 0), (ia, ib)
    a.step(1)
    b.step(1)
    fa = a.field().reshape(-1, S)
    fb = b.field().reshape(-1, S)
    keep = np.ones(len(fa), dtype=bool)
    keep[j] = False
    assert np.array_equal(fa[keep], fb[keep])
    assert not np.array_equal(fa[j], fb[j])


@pytest.mark.parametrize("grid,S,kernels", [((100, 144, 170), 8, [6, 5, 1]), ((170, 190, 100), 8, [6, 5, 1]),
                                            ((170, 100, 190), 8, [6, 1, 1])])
def test_two_lane_mid_lines_past_one_round(grid, S, kernels):
    """y lines of 137..208 voxels whose count needs more than one round of
    the register file (256 CUs) at one lane per line take two lanes per line;
    z lines of that length (the last sweep) keep one lane per line, as do
    lines within one round; one step matches the oracle"""
    rng = np.random.default_rng(sum(grid) + S)
    p = orc.Problem(3, grid, (20, 20, 20), 0.01, 10 ** rng.uniform(0, 5, S), rng.uniform(0.01, 1, S))
    f0 = rng.uniform(0, 1, p.shape)
    d = gpu_domain(p, field=f0)
    assert d.sweep_plan()["kernels"] == kernels
    want = p.diffuse(f0.astype(np.float32).astype(np.float64))
    d.step(1)
    assert rel_linf(d.field(), want).max() < 2e-6


@pytest.mark.parametrize("dims,grid,last", [(3, (6, 5, 240), 2), (2, (7, 240, 1), 1)])
def test_last_sweep_kernel_follows_the_bulk(dims, grid, last):
    """Last-axis lines of 209..256 voxels take two lanes per line.  z lines
    keep them with a uniform bulk riding on their stores; the y lines of a 2-D
    domain go to one lane per line (which applies it) when the bulk list has
    that shape (one run per line, an eighth of the voxels), whatever its
    rates.  The kernel follows the list, and the steps match the oracle"""
    from physicore_amd import bfvm
    from physicore_amd import workloads as wl

    w, wo, _ = _uniform_bulk_workloads(dims, grid, 2, seed=5, shape="dense")
    wn = dataclasses.replace(w, bulk=None)
    d = wl.make_domain(wn)
    assert d.sweep_plan()["kernels"][last] == 5
    d.step(2)
    p, field, dirs, bulk, ag = sc.oracle_inputs(wn)
    s = orc.Solver(p, field.astype(np.float32).astype(np.float64), dirichlet=dirs, bulk=bulk, agents=ag)
    s.solve(2)
    assert rel_linf(d.field(), s.field).max() < 1e-5
    with_bulk = 5 if last == 2 else 1
    d.set_bulk(*w.bulk)
    assert d.sweep_plan()["kernels"][last] == with_bulk
    d.set_bulk(*wo.bulk)  # rates that differ at one voxel: the same kernel
    assert d.sweep_plan()["kernels"][last] == with_bulk
    vox, sup, upt, tgt = w.bulk
    d.set_bulk(vox[:20], sup[:20], upt[:20], tgt[:20])  # a few voxels: two lanes
    assert d.sweep_plan()["kernels"][last] == 5
    d.set_bulk(np.zeros(0, dtype=np.int64), np.zeros((0, 2)), np.zeros((0, 2)), np.zeros((0, 2)))
    assert d.sweep_plan()["kernels"][last] == 5
    # the uniform bulk applied by the last sweep (one stage launch fewer in a
    # batch) against the oracle
    d2 = wl.make_domain(w)
    assert d2.sweep_plan()["kernels"][last] == with_bulk
    d3 = wl.make_domain(wo)
    b2, b3 = bfvm.Batch([d2]), bfvm.Batch([d3])
    assert b2.info()["stage_launches"] == b3.info()["stage_launches"] - 1
    b2.close()
    b3.close()
    d2.step(3)
    p, field, dirs, bulk, ag = sc.oracle_inputs(w)
    s = orc.Solver(p, field.astype(np.float32).astype(np.float64), dirichlet=dirs, bulk=bulk, agents=ag)
    s.solve(3)
    assert rel_linf(d2.field(), s.field).max() < 1e-5


@pytest.mark.parametrize("shape", ["dense", "ball"])
@pytest.mark.parametrize("dims,grid,S", [(3, (9, 7, 12), 3), (3, (24, 20, 40), 4), (2, (13, 11, 1), 2)])
def test_uniform_bulk_in_last_sweep_matches_oracle(dims, grid, S, shape):
    """Five steps with the bulk applied by the last sweep against the oracle
    (bulk, faces and cells) at the multi-step bar; a batch of two such domains
    equals each stepped alone bit for bit."""
    from physicore_amd import bfvm
    from physicore_amd import workloads as wl

    w, _, _ = _uniform_bulk_workloads(dims, grid, S, seed=99 + S, shape=shape)
    a = wl.make_domain(w, reduce="sorted")
    a.step(w.steps)
    pair = [wl.make_domain(w, reduce="sorted") for _ in range(2)]
    bt = bfvm.Batch(pair)
    bt.step(w.steps)
    bt.sync()
    bt.close()
    for d in pair:
        assert np.array_equal(d.field(), a.field())
    p, field, dirs, bulk, ag = sc.oracle_inputs(w)
    s = orc.Solver(p, field.astype(np.float32).astype(np.float64), dirichlet=dirs, bulk=bulk, agents=ag)
    s.solve(w.steps)
    err = rel_linf(a.field(), s.field)
    assert err.max() < 1e-5, err
