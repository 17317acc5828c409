"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke() entry point and by
bench.py's cpu_baseline leg, as the checker.  The product never imports it.

The oracle restates the reference OpenMP BioFVM solver
(reactions-diffusion/biofvm/kernels/openmp_solver/src/*.cpp); see bfvm_oracle.h
for the per-function citations.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

i64p = C.POINTER(C.c_int64)
f64p = C.POINTER(C.c_double)
u8p = C.POINTER(C.c_uint8)


class Domain(C.Structure):
    _fields_ = [("dims", C.c_int32), ("substrates", C.c_int32),
                ("nx", C.c_int64), ("ny", C.c_int64), ("nz", C.c_int64),
                ("dx", C.c_int64), ("dy", C.c_int64), ("dz", C.c_int64),
                ("bbox_min", C.c_int64 * 3), ("dt", C.c_double),
                ("diffusion", f64p), ("decay", f64p), ("openmp_quirks", C.c_int32)]


class Dirichlet(C.Structure):
    _fields_ = [("min_values", f64p * 3), ("min_conds", u8p * 3),
                ("max_values", f64p * 3), ("max_conds", u8p * 3),
                ("interior_count", C.c_int64), ("interior_voxels", i64p),
                ("interior_values", f64p), ("interior_conds", u8p)]


class Bulk(C.Structure):
    _fields_ = [("supply", f64p), ("uptake", f64p), ("target", f64p)]


class Agents(C.Structure):
    _fields_ = [("count", C.c_int64), ("positions", f64p), ("secretion", f64p),
                ("saturation", f64p), ("uptake", f64p), ("net_export", f64p),
                ("internalized", f64p), ("fraction_released", f64p),
                ("volumes", f64p), ("is_active", u8p)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        L = C.CDLL(path)
        L.orc_factors.argtypes = [C.POINTER(Domain), C.c_int, f64p, f64p, f64p]
        L.orc_coef_substrate.argtypes = [C.POINTER(Domain), C.c_int64]
        L.orc_coef_substrate.restype = C.c_int64
        L.orc_diffuse_f64.argtypes = [C.POINTER(Domain), f64p, C.c_int64]
        L.orc_diffuse_f32.argtypes = [C.POINTER(Domain), C.POINTER(C.c_float), C.c_int64]
        L.orc_sweep_f64.argtypes = [C.POINTER(Domain), f64p, C.c_int]
        L.orc_dirichlet_f64.argtypes = [C.POINTER(Domain), f64p, C.POINTER(Dirichlet)]
        L.orc_bulk_f64.argtypes = [C.POINTER(Domain), f64p, C.POINTER(Bulk)]
        L.orc_solver_create.argtypes = [C.POINTER(Domain), f64p, C.POINTER(Dirichlet),
                                        C.POINTER(Bulk), C.POINTER(Agents), C.c_int]
        L.orc_solver_create.restype = C.c_void_p
        L.orc_solver_create_f32.argtypes = [C.POINTER(Domain), C.POINTER(C.c_float),
                                            C.POINTER(Dirichlet), C.POINTER(Bulk),
                                            C.POINTER(Agents), C.c_int]
        L.orc_solver_create_f32.restype = C.c_void_p
        L.orc_solver_solve.argtypes = [C.c_void_p, C.c_int64]
        L.orc_solver_recompute.argtypes = [C.c_void_p]
        L.orc_solver_release.argtypes = [C.c_void_p, C.c_int64]
        L.orc_solver_reduced.argtypes = [C.c_void_p, f64p, f64p, f64p, i64p,
                                         C.POINTER(C.c_int32)]
        L.orc_solver_destroy.argtypes = [C.c_void_p]
        L.orc_num_threads.restype = C.c_int
        _LIB = L
    return _LIB


def _p(a, t=f64p):
    return None if a is None else a.ctypes.data_as(t)


def grid_shape(dims, bmin, bmax, vshape):
    """cartesian_mesh grid shape: ceil((max-min)/d) per axis below dims, else 1
    (reactions-diffusion/biofvm/src/mesh.cpp:8-28)."""
    g = [1, 1, 1]
    for a in range(dims):
        g[a] = (bmax[a] - bmin[a] + vshape[a] - 1) // vshape[a]
    return g


class Problem:
    """One BioFVM domain in the oracle's terms; keeps every array alive."""

    def __init__(self, dims, grid, voxel, dt, diffusion, decay, bbox_min=(0, 0, 0),
                 openmp_quirks=False):
        self.dims = dims
        self.nx, self.ny, self.nz = (int(g) for g in grid)
        self.voxel = tuple(int(v) for v in voxel)
        self.S = len(diffusion)
        self.dt = float(dt)
        self.diffusion = np.ascontiguousarray(diffusion, dtype=np.float64)
        self.decay = np.ascontiguousarray(decay, dtype=np.float64)
        self.bbox_min = tuple(int(b) for b in bbox_min)
        self.openmp_quirks = bool(openmp_quirks)
        self.dom = Domain(dims, self.S, self.nx, self.ny, self.nz, *self.voxel,
                          (C.c_int64 * 3)(*self.bbox_min), self.dt,
                          _p(self.diffusion), _p(self.decay), int(self.openmp_quirks))

    @property
    def shape(self):
        return (self.nz, self.ny, self.nx, self.S)

    def factors(self, axis):
        n = (self.nx, self.ny, self.nz)[axis]
        b = np.zeros(n * self.S)
        e = np.zeros(max(n - 1, 1) * self.S)
        c = np.zeros(self.S)
        lib().orc_factors(C.byref(self.dom), axis, _p(b), _p(e), _p(c))
        return b.reshape(n, self.S), e[: (n - 1) * self.S].reshape(max(n - 1, 0), self.S), c

    def coef_substrate(self, column):
        return lib().orc_coef_substrate(C.byref(self.dom), int(column))

    def diffuse(self, field, iterations=1):
        assert field.flags.c_contiguous and field.size == self.nx * self.ny * self.nz * self.S
        if field.dtype == np.float64:
            lib().orc_diffuse_f64(C.byref(self.dom), _p(field), iterations)
        elif field.dtype == np.float32:
            lib().orc_diffuse_f32(C.byref(self.dom), field.ctypes.data_as(C.POINTER(C.c_float)),
                                  iterations)
        else:
            raise TypeError(field.dtype)
        return field

    def sweep(self, field, axis):
        assert field.dtype == np.float64 and field.flags.c_contiguous
        lib().orc_sweep_f64(C.byref(self.dom), _p(field), axis)
        return field


class DirichletSpec:
    """Face and interior Dirichlet conditions (microenvironment.h:65-75)."""

    def __init__(self, S, dims):
        self.S, self.dims = S, dims
        self.min_values = [None] * 3
        self.min_conds = [None] * 3
        self.max_values = [None] * 3
        self.max_conds = [None] * 3
        self.interior_voxels = np.zeros((0, dims), dtype=np.int64)
        self.interior_values = np.zeros((0, S))
        self.interior_conds = np.zeros((0, S), dtype=np.uint8)

    def set_face(self, axis, side, values, conds):
        v = np.ascontiguousarray(values, dtype=np.float64)
        c = np.ascontiguousarray(conds, dtype=np.uint8)
        if side == "min":
            self.min_values[axis], self.min_conds[axis] = v, c
        else:
            self.max_values[axis], self.max_conds[axis] = v, c

    def set_interior(self, voxels, values, conds):
        self.interior_voxels = np.ascontiguousarray(voxels, dtype=np.int64).reshape(-1, self.dims)
        self.interior_values = np.ascontiguousarray(values, dtype=np.float64).reshape(-1, self.S)
        self.interior_conds = np.ascontiguousarray(conds, dtype=np.uint8).reshape(-1, self.S)

    def struct(self):
        d = Dirichlet()
        for a in range(3):
            d.min_values[a] = _p(self.min_values[a])
            d.min_conds[a] = _p(self.min_conds[a], u8p)
            d.max_values[a] = _p(self.max_values[a])
            d.max_conds[a] = _p(self.max_conds[a], u8p)
        d.interior_count = len(self.interior_voxels)
        d.interior_voxels = _p(self.interior_voxels, i64p)
        d.interior_values = _p(self.interior_values)
        d.interior_conds = _p(self.interior_conds, u8p)
        return d


class AgentSpec:
    """Agent SoA (agent_data_generic_storage.h:16-31), FP64 like the reference."""

    def __init__(self, n, S, dims):
        self.n, self.S, self.dims = n, S, dims
        self.positions = np.zeros((n, dims))
        self.secretion = np.zeros((n, S))
        self.saturation = np.zeros((n, S))
        self.uptake = np.zeros((n, S))
        self.net_export = np.zeros((n, S))
        self.internalized = np.zeros((n, S))
        self.fraction_released = np.zeros((n, S))
        self.volumes = np.zeros(n)
        self.is_active = np.ones(n, dtype=np.uint8)

    def struct(self):
        for k in ("positions", "secretion", "saturation", "uptake", "net_export",
                  "internalized", "fraction_released", "volumes", "is_active"):
            a = getattr(self, k)
            if not a.flags.c_contiguous:
                setattr(self, k, np.ascontiguousarray(a))
        return Agents(self.n, _p(self.positions), _p(self.secretion), _p(self.saturation),
                      _p(self.uptake), _p(self.net_export), _p(self.internalized),
                      _p(self.fraction_released), _p(self.volumes), _p(self.is_active, u8p))


def dirichlet(problem, field, spec):
    s = spec.struct()
    lib().orc_dirichlet_f64(C.byref(problem.dom), _p(field), C.byref(s))
    return field


def bulk(problem, field, supply, uptake, target):
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (supply, uptake, target)]
    b = Bulk(*(_p(a) for a in arrs))
    lib().orc_bulk_f64(C.byref(problem.dom), _p(field), C.byref(b))
    return field


class Solver:
    """Full four-stage solver (openmp_solver.cpp:23-40).  All inputs are borrowed
    and read live on every step, like the OpenMP backend."""

    def __init__(self, problem, field, dirichlet=None, bulk=None, agents=None,
                 compute_internalized=False):
        self.problem = problem
        self.field = field
        self._dir = dirichlet.struct() if dirichlet is not None else None
        self._bulk_arrays = None
        if bulk is not None:
            self._bulk_arrays = [np.ascontiguousarray(a, dtype=np.float64) for a in bulk]
            self._bulk = Bulk(*(_p(a) for a in self._bulk_arrays))
        else:
            self._bulk = None
        self.agents = agents
        self._ag = agents.struct() if agents is not None else None
        L = lib()
        args = (C.byref(problem.dom),)
        dirp = C.byref(self._dir) if self._dir is not None else None
        bulkp = C.byref(self._bulk) if self._bulk is not None else None
        agp = C.byref(self._ag) if self._ag is not None else None
        if field.dtype == np.float64:
            self.h = L.orc_solver_create(*args, _p(field), dirp, bulkp, agp, int(compute_internalized))
        else:
            self.h = L.orc_solver_create_f32(*args, field.ctypes.data_as(C.POINTER(C.c_float)),
                                             dirp, bulkp, agp, int(compute_internalized))

    def solve(self, iterations=1):
        rc = lib().orc_solver_solve(self.h, iterations)
        if rc != 0:
            raise RuntimeError("oracle: an active cell lies outside the grid")

    def recompute(self):
        lib().orc_solver_recompute(self.h)

    def release(self, index):
        if lib().orc_solver_release(self.h, index) != 0:
            raise RuntimeError("oracle: bad agent index")

    def reduced(self):
        n, S = (self.agents.n, self.problem.S) if self.agents is not None else (0, self.problem.S)
        V = self.problem.nx * self.problem.ny * self.problem.nz
        num = np.zeros(n * S)
        den = np.zeros(n * S)
        fac = np.zeros(n * S)
        ballots = np.zeros(V, dtype=np.int64)
        conflict = C.c_int32(0)
        lib().orc_solver_reduced(self.h, _p(num), _p(den), _p(fac), _p(ballots, i64p),
                                 C.byref(conflict))
        return num.reshape(n, S), den.reshape(n, S), fac.reshape(n, S), ballots, bool(conflict.value)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_solver_destroy(self.h)
            self.h = None


def num_threads():
    return lib().orc_num_threads()
