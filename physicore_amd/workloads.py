"""Synthetic inputs of the benchmark workloads, as plain numpy arrays.

Each generator returns a Workload: one microenvironment (mesh, coefficients,
initial field, Dirichlet faces and interior voxels, a sparse bulk list and the
agent arrays) plus the step count of the workload.  The benchmark and the
tests build their GPU domains from the same Workload (make_domain), and the
tests build the oracle's inputs from it too, so both sides start from
identical arrays.  Seeds are fixed per workload (numpy default_rng).

  c1()         50^3 x 1 substrate, Dirichlet 38 on all faces, 1000 uptaking cells, 100 steps
  c2()         200^3 x 4 substrates, x-min Dirichlet, bulk sphere r=500um, 1e5 cells, 1000 steps
  c3_domain()  one domain of the 64-domain mixed ensemble (sides {32..160}, S 1-8, 0-50k cells), 500 steps
  c4()         400x400x100 x 2 substrates at dx=10, 2e6 cells in a Gaussian spheroid, 200 steps
  reference_benchmark()
               the reference's own benchmark program
               (reactions-diffusion/biofvm/kernels/openmp_solver/examples/benchmark.cpp:74-204):
               250^3 x 4, bulk in every voxel, Dirichlet on all six faces,
               2e6 lattice agents plus one sharing agent, internalized tracking,
               recompute every 10 steps, 100 steps

The served stream (Poisson arrivals) is physicore_amd.serve.poisson_stream.
"""
import dataclasses
from typing import Optional

import numpy as np


@dataclasses.dataclass
class Workload:
    name: str
    dims: int
    grid: tuple                               # (nx, ny, nz)
    voxel: tuple                              # (dx, dy, dz)
    dt: float
    diffusion: np.ndarray                     # [S]
    decay: np.ndarray                         # [S]
    steps: int
    initial: Optional[np.ndarray] = None      # [S], the field when `field` is None
    field: Optional[np.ndarray] = None        # (nz, ny, nx, S) FP64
    faces: Optional[tuple] = None             # (values[6, S], conds[6, S]); x-min, x-max, y-min, ...
    interior: Optional[tuple] = None          # (voxels[k, dims], values[k, S], conds[k, S])
    bulk: Optional[tuple] = None              # (voxels[k] linear, supply[k, S], uptake[k, S], target[k, S])
    agents: Optional[dict] = None             # keyword arguments of bfvm.Domain.set_agents
    compute_internalized: bool = False
    recompute_every: int = 0                  # positional recompute period in steps (0: only at start)
    seed: Optional[int] = None

    @property
    def S(self):
        return len(self.diffusion)

    @property
    def voxels(self):
        return int(np.prod(self.grid))

    @property
    def units(self):
        """voxel*substrate values advanced per step"""
        return self.voxels * self.S

    @property
    def n_agents(self):
        return 0 if self.agents is None else len(self.agents["volumes"])

    def field_or_initial(self):
        if self.field is not None:
            return self.field
        f = np.empty((self.grid[2], self.grid[1], self.grid[0], self.S))
        f[...] = np.zeros(self.S) if self.initial is None else self.initial
        return f


def agent_arrays(n, S, dims, positions):
    """Agent SoA with the reference's defaults (agent_data_generic_storage.h:49-62:
    rates, volumes and internalized amounts 0, active)."""
    return dict(positions=np.ascontiguousarray(positions[:, :dims], dtype=np.float64),
                secretion=np.zeros((n, S)), saturation=np.zeros((n, S)), uptake=np.zeros((n, S)),
                net_export=np.zeros((n, S)), internalized=None, fraction_released=np.zeros((n, S)),
                volumes=np.zeros(n), is_active=np.ones(n, dtype=np.uint8))


def x_min_faces(S, value=1.0):
    vals = np.zeros((6, S))
    conds = np.zeros((6, S), dtype=np.uint8)
    vals[0], conds[0] = value, 1
    return vals, conds


def _d_lambda(rng, S):
    D = 10.0 ** rng.uniform(0.0, 5.0, S)
    lam = rng.uniform(0.01, 1.0, S)
    return D, lam


def _uniform_cells(rng, n, S, box):
    pos = rng.uniform(0.0, box, (n, 3))
    np.minimum(pos, np.nextafter(np.asarray(box, dtype=np.float64), 0), out=pos)
    a = agent_arrays(n, S, 3, pos)
    a["secretion"][:] = rng.uniform(0.0, 10.0, (n, S))
    a["uptake"][:] = rng.uniform(0.0, 10.0, (n, S))
    a["saturation"][:] = 1.0
    a["volumes"][:] = 2494.0
    return a


def c1(seed=42):
    """50^3 voxels (1000um cube, dx=20), S=1, D=1000, decay 0.1, dt=0.01, field 0,
    Dirichlet 38 on all six faces, 1000 cells uniform (volume 2494, secretion 0,
    uptake 10, saturation 38); 100 steps."""
    rng = np.random.default_rng(seed)
    n = 1000
    a = agent_arrays(n, 1, 3, rng.uniform(0.0, 1000.0, (n, 3)))
    a["uptake"][:] = 10.0
    a["saturation"][:] = 38.0
    a["volumes"][:] = 2494.0
    vals = np.full((6, 1), 38.0)
    conds = np.ones((6, 1), dtype=np.uint8)
    return Workload("50^3x1, Dirichlet 38 on all faces, 1000 cells", 3, (50, 50, 50), (20, 20, 20), 0.01,
                    np.array([1000.0]), np.array([0.1]), 100, initial=np.zeros(1), faces=(vals, conds),
                    agents=a, seed=seed)


def c2(side=200, S=4, n_cells=100_000, seed=43, with_cells=True, with_bulk=True):
    """side^3 voxels at dx=20, S substrates with D log-uniform in [1, 1e5] and decay
    uniform in [0.01, 1], field U[0,1], Dirichlet 1.0 on the x-min face, bulk supply
    0.5/min (target 1, uptake 0) inside a central sphere of r=500um, uniform cells
    with secretion/uptake U[0,10] (saturation 1, volume 2494); 1000 steps."""
    rng = np.random.default_rng(seed)
    dx = 20
    box = side * dx
    D, lam = _d_lambda(rng, S)
    field = rng.uniform(0.0, 1.0, (side, side, side, S))
    bulk = None
    if with_bulk:
        c = (np.arange(side) + 0.5) * dx - box / 2.0
        zz, yy, xx = np.meshgrid(c, c, c, indexing="ij")
        vox = np.nonzero(((xx ** 2 + yy ** 2 + zz ** 2) <= 500.0 ** 2).reshape(-1))[0]
        k = len(vox)
        bulk = (vox, np.full((k, S), 0.5), np.zeros((k, S)), np.ones((k, S)))
    agents = _uniform_cells(rng, n_cells, S, box) if with_cells and n_cells > 0 else None
    return Workload(f"{side}^3x{S}, x-min Dirichlet, bulk sphere r=500um, {n_cells if agents else 0} cells",
                    3, (side, side, side), (dx, dx, dx), 0.01, D, lam, 1000, field=field,
                    faces=x_min_faces(S), bulk=bulk, agents=agents, seed=seed)


C3_SIDES = tuple(range(32, 161, 16))


def c3_domain(k, seed=44, sides=C3_SIDES, max_S=8, max_cells=50_000, with_field=True):
    """Domain k of the mixed ensemble: each axis side drawn from `sides`
    (anisotropic), dx=20, S uniform in 1..max_S with D/decay as c2, field U[0,1],
    0..max_cells uniform cells (secretion/uptake U[0,10], saturation 1, volume
    2494); 500 steps.  Domain k has its own generator (seeded by (seed, k)), so
    any domain can be rebuilt alone."""
    rng = np.random.default_rng([seed, k])
    n = tuple(int(rng.choice(sides)) for _ in range(3))
    S = int(rng.integers(1, max_S + 1))
    D, lam = _d_lambda(rng, S)
    field = rng.uniform(0.0, 1.0, (n[2], n[1], n[0], S)) if with_field else None
    m = int(rng.integers(0, max_cells + 1))
    agents = _uniform_cells(rng, m, S, np.array(n, dtype=np.float64) * 20.0) if m else None
    return Workload(f"ensemble domain {k}: {n[0]}x{n[1]}x{n[2]}x{S}, {m} cells", 3, n, (20, 20, 20), 0.01, D,
                    lam, 500, field=field, initial=np.zeros(S), agents=agents, seed=seed)


def c3(seed=44, count=64):
    """The 64-domain ensemble, one domain at a time (a generator)."""
    for k in range(count):
        yield c3_domain(k, seed)


def c4(seed=45, n_cells=2_000_000, nx=400, ny=400, nz=100):
    """nx x ny x nz voxels at dx=10 (400x400x100: box 4000x4000x1000), S=2 with
    D={1e5, 100}, decay {0.1, 0.01}, field U[0,1]; cells drawn from an isotropic
    Gaussian (sigma 300um) at the box centre, redrawn until inside the box;
    secretion/uptake U[0,10] (saturation 1, volume 2494), net export U[-1,1];
    200 steps."""
    rng = np.random.default_rng(seed)
    dx = 10
    field = rng.uniform(0.0, 1.0, (nz, ny, nx, 2))
    box = np.array([nx * dx, ny * dx, nz * dx], dtype=np.float64)
    pos = np.empty((0, 3))
    while len(pos) < n_cells:
        c = rng.normal(box / 2, 300.0, (n_cells, 3))
        c = c[((c >= 0) & (c < box)).all(axis=1)]
        pos = np.concatenate([pos, c])[:n_cells]
    a = agent_arrays(n_cells, 2, 3, pos)
    a["secretion"][:] = rng.uniform(0.0, 10.0, (n_cells, 2))
    a["uptake"][:] = rng.uniform(0.0, 10.0, (n_cells, 2))
    a["saturation"][:] = 1.0
    a["volumes"][:] = 2494.0
    a["net_export"][:] = rng.uniform(-1.0, 1.0, (n_cells, 2))
    return Workload(f"{nx}x{ny}x{nz}x2 at dx=10, {n_cells} Gaussian cells", 3, (nx, ny, nz), (dx, dx, dx), 0.01,
                    np.array([1e5, 100.0]), np.array([0.1, 0.01]), 200, field=field, agents=a, seed=seed)


def reference_benchmark(side_um=5000, n_agents=2_000_000, steps=100):
    """The reference's benchmark program (openmp_solver/examples/benchmark.cpp):
    mesh [0, side_um)^3 at dx=20 (250^3 voxels), 4 substrates with initial 1000,
    D 10000, decay 0.5, dt 0.01, internalized tracking on (:77-90); a bulk
    functor with supply = target = uptake = 0.01*(s+1) in every voxel (:96-113);
    Dirichlet on all six faces, 0 on the min faces and 0.5*(axis+1) on the max
    faces (:116-131); n_agents agents on a 20um lattice from the origin, x
    fastest, plus one more agent at the origin (:17-50, :133), with the
    agent container's default values (all rates, volumes and internalized
    amounts 0); the positional data recomputed every 10 steps (:169)."""
    S, dx = 4, 20
    n = (side_um + dx - 1) // dx
    V = n ** 3
    vox = np.arange(V, dtype=np.int64)
    coef = np.broadcast_to(0.01 * (np.arange(S) + 1.0), (V, S))
    bulk = (vox, coef, coef, coef)
    vals = np.zeros((6, S))
    for a in range(3):
        vals[2 * a + 1] = 0.5 * (a + 1)
    conds = np.ones((6, S), dtype=np.uint8)
    # lattice walk of make_agents: x += 20, wrapping into y, then z
    i = np.arange(n_agents, dtype=np.int64)
    per_row = (side_um + dx - 1) // dx
    x = (i % per_row) * dx
    y = ((i // per_row) % per_row) * dx
    z = (i // (per_row * per_row)) * dx
    pos = np.stack([x, y, z], axis=1).astype(np.float64)
    pos = np.concatenate([pos, np.zeros((1, 3))])
    a = agent_arrays(n_agents + 1, S, 3, pos)
    return Workload(f"reference benchmark: {n}^3x{S}, bulk in every voxel, Dirichlet on all faces, "
                    f"{n_agents}+1 lattice agents, internalized, recompute every 10 steps", 3, (n, n, n),
                    (dx, dx, dx), 0.01, np.full(S, 10000.0), np.full(S, 0.5), steps,
                    initial=np.full(S, 1000.0), faces=(vals, conds), bulk=bulk, agents=a,
                    compute_internalized=True, recompute_every=10)


def make_domain(w, precision="fp64", reduce="atomic", openmp_quirks=False, upload_field=True):
    """A bfvm.Domain holding workload `w` (field, hooks and agents uploaded)."""
    from physicore_amd import bfvm

    d = bfvm.Domain(w.dims, w.grid, w.voxel, w.dt, w.diffusion, w.decay,
                    initial=w.initial if w.initial is not None else np.zeros(w.S),
                    compute_internalized=w.compute_internalized, precision=precision, reduce=reduce,
                    openmp_quirks=openmp_quirks)
    if upload_field and w.field is not None:
        d.upload_field(w.field)
    if w.faces is not None:
        d.set_dirichlet_faces(*w.faces)
    if w.interior is not None:
        d.set_dirichlet_interior(*w.interior)
    if w.bulk is not None:
        d.set_bulk(*w.bulk)
    if w.agents is not None:
        d.set_agents(**w.agents)
    return d
