"""ctypes binding of libbfvm_hip.so (include/bfvm_hip.h).

The HIP library is the only compute path: there is no CPU fallback, and every
call raises BfvmError when the library is missing or a call fails.
"""
import ctypes as C
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "lib", "libbfvm_hip.so")
HEADER = os.path.join(ROOT, "include", "bfvm_hip.h")

OK = 0
STATUS = {0: "OK", -1: "INVALID", -2: "HIP", -3: "NO_MEMORY", -4: "OUT_OF_GRID",
          -5: "UNSUPPORTED", -6: "STATE"}
PRECISION = {"fp64": 0, "fp32": 1}
REDUCE = {"atomic": 0, "sorted": 1}

f64p = C.POINTER(C.c_double)
f32p = C.POINTER(C.c_float)
i64p = C.POINTER(C.c_int64)
u8p = C.POINTER(C.c_uint8)


class BfvmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"bfvm {STATUS.get(code, code)}: {msg}")
        self.code = code


class DomainDesc(C.Structure):
    _fields_ = [("dims", C.c_int32), ("substrates", C.c_int32),
                ("grid", C.c_int64 * 3), ("voxel", C.c_int64 * 3), ("bbox_min", C.c_int64 * 3),
                ("dt", C.c_double), ("diffusion_coefficients", f64p), ("decay_rates", f64p),
                ("initial_conditions", f64p), ("compute_internalized", C.c_int32),
                ("precision", C.c_int32), ("reduce_mode", C.c_int32), ("openmp_quirks", C.c_int32)]


class Tuning(C.Structure):
    _fields_ = [("persistent_wgs_per_cu", C.c_int32), ("force_lds_sweeps", C.c_int32),
                ("x_sweep", C.c_int32), ("pipeline", C.c_int32), ("round_nearest", C.c_int32),
                ("lines_per_block", C.c_int32), ("split_xy", C.c_int32),
                ("reserved", C.c_int32 * 1)]


_LIB = None

_SIGS = {
    "bfvm_last_error": ([C.c_void_p], C.c_char_p),
    "bfvm_device_count": ([], C.c_int),
    "bfvm_create": ([C.POINTER(DomainDesc), C.POINTER(C.c_void_p)], C.c_int),
    "bfvm_destroy": ([C.c_void_p], None),
    "bfvm_set_dirichlet_faces": ([C.c_void_p, f64p, u8p], C.c_int),
    "bfvm_set_dirichlet_interior": ([C.c_void_p, C.c_int64, i64p, f64p, u8p], C.c_int),
    "bfvm_set_bulk": ([C.c_void_p, C.c_int64, i64p, f64p, f64p, f64p], C.c_int),
    "bfvm_set_agents": ([C.c_void_p, C.c_int64] + [f64p] * 8 + [u8p], C.c_int),
    "bfvm_recompute_positional_data": ([C.c_void_p], C.c_int),
    "bfvm_agent_positions_device_ptr": ([C.c_void_p, C.POINTER(C.c_void_p)], C.c_int),
    "bfvm_step": ([C.c_void_p, C.c_int64], C.c_int),
    "bfvm_step_graph": ([C.c_void_p, C.c_int64, C.c_int64], C.c_int),
    "bfvm_sync": ([C.c_void_p], C.c_int),
    "bfvm_download_field_f64": ([C.c_void_p, f64p], C.c_int),
    "bfvm_upload_field_f64": ([C.c_void_p, f64p], C.c_int),
    "bfvm_download_field_f32": ([C.c_void_p, f32p], C.c_int),
    "bfvm_upload_field_f32": ([C.c_void_p, f32p], C.c_int),
    "bfvm_snapshot_begin": ([C.c_void_p, f32p, C.POINTER(C.c_int64)], C.c_int),
    "bfvm_snapshot_wait": ([C.c_void_p, C.c_int64], C.c_int),
    "bfvm_host_alloc": ([C.c_size_t, C.POINTER(C.c_void_p)], C.c_int),
    "bfvm_host_free": ([C.c_void_p], C.c_int),
    "bfvm_download_internalized": ([C.c_void_p, f64p], C.c_int),
    "bfvm_release_internalized": ([C.c_void_p, C.c_int64], C.c_int),
    "bfvm_total_mass": ([C.c_void_p, f64p], C.c_int),
    "bfvm_total_mass_static": ([C.c_void_p, f64p], C.c_int),
    "bfvm_cell_sums": ([C.c_void_p, i64p, i64p, f64p, f64p, f64p, C.POINTER(C.c_int32)], C.c_int),
    "bfvm_field_device_ptr": ([C.c_void_p, C.POINTER(C.c_void_p)], C.c_int),
    "bfvm_stream": ([C.c_void_p, C.POINTER(C.c_void_p)], C.c_int),
    "bfvm_time_stages": ([C.c_void_p, C.c_int64, f64p], C.c_int),
    "bfvm_time_sweep": ([C.c_void_p, C.c_int32, C.c_int64, f64p], C.c_int),
    "bfvm_get_tuning": ([C.c_void_p, C.POINTER(Tuning)], C.c_int),
    "bfvm_sweep_plan": ([C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)], C.c_int),
    "bfvm_set_tuning": ([C.c_void_p, C.POINTER(Tuning)], C.c_int),
    "bfvm_batch_create": ([C.POINTER(C.c_void_p), C.c_int32, C.POINTER(C.c_void_p)], C.c_int),
    "bfvm_batch_destroy": ([C.c_void_p], None),
    "bfvm_batch_last_error": ([C.c_void_p], C.c_char_p),
    "bfvm_batch_step": ([C.c_void_p, C.c_int64, C.c_int64], C.c_int),
    "bfvm_batch_sync": ([C.c_void_p], C.c_int),
    "bfvm_batch_info": ([C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)], C.c_int),
    "bfvm_batch_mark": ([C.c_void_p, C.POINTER(C.c_int64)], C.c_int),
    "bfvm_batch_wait": ([C.c_void_p, C.c_int64], C.c_int),
    "bfvm_batch_query": ([C.c_void_p, C.c_int64, C.POINTER(C.c_int32)], C.c_int),
    "bfvm_batch_autotune": ([C.c_void_p, C.c_int64, C.POINTER(Tuning)], C.c_int),
}


def declared_symbols(header=HEADER):
    """Function names declared in include/bfvm_hip.h."""
    text = open(header).read()
    return sorted(set(re.findall(r"\b(bfvm_[a-z0-9_]+)\s*\(", text)))


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise BfvmError(-6, f"{LIB_PATH} is not built; run `python -m physicore_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _LIB = L
    return _LIB


def device_count():
    return lib().bfvm_device_count()


def _p(a, t=f64p):
    return None if a is None else a.ctypes.data_as(t)


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a if shape is None else a.reshape(shape)


class PinnedField:
    """Pinned host memory (bfvm_host_alloc) shaped like a Domain's FP32 field,
    the destination of Domain.snapshot_begin."""

    def __init__(self, shape):
        self.shape = tuple(shape)
        n = int(np.prod(self.shape))
        self._ptr = C.c_void_p()
        rc = lib().bfvm_host_alloc(n * 4, C.byref(self._ptr))
        if rc != OK:
            raise BfvmError(rc, lib().bfvm_last_error(None).decode())
        buf = (C.c_float * n).from_address(self._ptr.value)
        self.array = np.ctypeslib.as_array(buf).reshape(self.shape)

    def close(self):
        if getattr(self, "_ptr", None) and self._ptr.value:
            self.array = None
            lib().bfvm_host_free(self._ptr)
            self._ptr = C.c_void_p()

    def __del__(self):
        self.close()


class Domain:
    """One microenvironment resident on the GPU (a bfvm_ctx)."""

    def __init__(self, dims, grid, voxel, dt, diffusion, decay, initial=None, bbox_min=(0, 0, 0),
                 compute_internalized=False, precision="fp64", reduce="atomic", openmp_quirks=False):
        self.dims = int(dims)
        self.grid = tuple(int(g) for g in grid)
        self.voxel = tuple(int(v) for v in voxel)
        self.S = len(diffusion)
        self._D = _f64(diffusion)
        self._lam = _f64(decay)
        self._init = _f64(initial if initial is not None else np.zeros(self.S))
        desc = DomainDesc()
        desc.dims = self.dims
        desc.substrates = self.S
        for a in range(3):
            desc.grid[a] = self.grid[a]
            desc.voxel[a] = self.voxel[a]
            desc.bbox_min[a] = int(bbox_min[a])
        desc.dt = float(dt)
        desc.diffusion_coefficients = _p(self._D)
        desc.decay_rates = _p(self._lam)
        desc.initial_conditions = _p(self._init)
        desc.compute_internalized = int(bool(compute_internalized))
        desc.precision = PRECISION[precision]
        desc.reduce_mode = REDUCE[reduce]
        desc.openmp_quirks = int(bool(openmp_quirks))
        self._h = C.c_void_p()
        rc = lib().bfvm_create(C.byref(desc), C.byref(self._h))
        if rc != OK:
            raise BfvmError(rc, lib().bfvm_last_error(None).decode())
        self.n_agents = 0

    # ------------------------------------------------------------ plumbing
    def _check(self, rc):
        if rc != OK:
            raise BfvmError(rc, lib().bfvm_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().bfvm_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def shape(self):
        return (self.grid[2], self.grid[1], self.grid[0], self.S)

    # ---------------------------------------------------------------- hooks
    def set_dirichlet_faces(self, values, conds):
        """values/conds: [6, S], faces x-min, x-max, y-min, y-max, z-min, z-max."""
        v = _f64(values, (6 * self.S,))
        c = np.ascontiguousarray(conds, dtype=np.uint8).reshape(6 * self.S)
        self._check(lib().bfvm_set_dirichlet_faces(self._h, _p(v), _p(c, u8p)))

    def set_dirichlet_interior(self, voxels, values, conds):
        vox = np.ascontiguousarray(voxels, dtype=np.int64).reshape(-1)
        n = vox.size // self.dims
        v = _f64(values, (n * self.S,))
        c = np.ascontiguousarray(conds, dtype=np.uint8).reshape(n * self.S)
        self._check(lib().bfvm_set_dirichlet_interior(self._h, n, _p(vox, i64p), _p(v), _p(c, u8p)))

    def set_bulk(self, voxels, supply, uptake, target):
        vox = np.ascontiguousarray(voxels, dtype=np.int64).reshape(-1)
        n = vox.size
        arrs = [_f64(a, (n * self.S,)) for a in (supply, uptake, target)]
        self._check(lib().bfvm_set_bulk(self._h, n, _p(vox, i64p), *(_p(a) for a in arrs)))

    def set_agents(self, positions, secretion, saturation, uptake, net_export, internalized,
                   fraction_released, volumes, is_active):
        n = len(volumes)
        arrs = [_f64(a) for a in (positions, secretion, saturation, uptake, net_export)]
        internal = None if internalized is None else _f64(internalized)
        rest = [_f64(fraction_released), _f64(volumes)]
        act = np.ascontiguousarray(is_active, dtype=np.uint8)
        self._check(lib().bfvm_set_agents(self._h, n, *(_p(a) for a in arrs), _p(internal),
                                          *(_p(a) for a in rest), _p(act, u8p)))
        self.n_agents = n

    def recompute_positional_data(self):
        self._check(lib().bfvm_recompute_positional_data(self._h))

    def agent_positions_device_ptr(self):
        """Device address of the agent positions (n*dims FP64)."""
        p = C.c_void_p()
        self._check(lib().bfvm_agent_positions_device_ptr(self._h, C.byref(p)))
        return p.value

    # ----------------------------------------------------------------- step
    def step(self, iterations=1):
        self._check(lib().bfvm_step(self._h, int(iterations)))

    def step_graph(self, iterations, graph_steps=10):
        self._check(lib().bfvm_step_graph(self._h, int(iterations), int(graph_steps)))

    def sync(self):
        self._check(lib().bfvm_sync(self._h))

    # ------------------------------------------------------------ transfers
    def field(self, dtype=np.float32):
        if dtype == np.float32:
            out = np.empty(self.shape, dtype=np.float32)
            self._check(lib().bfvm_download_field_f32(self._h, _p(out, f32p)))
        else:
            out = np.empty(self.shape, dtype=np.float64)
            self._check(lib().bfvm_download_field_f64(self._h, _p(out)))
        return out

    def snapshot_begin(self, dst):
        """Queues a snapshot of the field (after the steps queued so far) into
        `dst`, a PinnedField; later steps run during its host copy.  Returns
        the ticket for snapshot_wait."""
        if tuple(dst.shape) != tuple(self.shape):
            raise ValueError("snapshot buffer shape mismatch")
        t = C.c_int64()
        self._check(lib().bfvm_snapshot_begin(self._h, _p(dst.array, f32p), C.byref(t)))
        return t.value

    def snapshot_wait(self, ticket):
        rc = lib().bfvm_snapshot_wait(self._h, int(ticket))
        if rc != OK:
            raise BfvmError(rc, lib().bfvm_last_error(None).decode())

    def upload_field(self, arr):
        a = np.ascontiguousarray(arr)
        if a.size != int(np.prod(self.shape)):
            raise ValueError("field size mismatch")
        if a.dtype == np.float32:
            self._check(lib().bfvm_upload_field_f32(self._h, _p(a, f32p)))
        else:
            a = _f64(a)
            self._check(lib().bfvm_upload_field_f64(self._h, _p(a)))

    def internalized(self):
        out = np.zeros((self.n_agents, self.S))
        self._check(lib().bfvm_download_internalized(self._h, _p(out)))
        return out

    def release_internalized(self, agent):
        self._check(lib().bfvm_release_internalized(self._h, int(agent)))

    def total_mass(self):
        out = np.zeros(self.S)
        self._check(lib().bfvm_total_mass(self._h, _p(out)))
        return out

    def total_mass_static(self):
        """total_mass through bfvm_total_mass_static (no device allocation)."""
        out = np.zeros(self.S)
        self._check(lib().bfvm_total_mass_static(self._h, _p(out)))
        return out

    def cell_sums(self):
        n = C.c_int64(0)
        conflict = C.c_int32(0)
        self._check(lib().bfvm_cell_sums(self._h, C.byref(n), None, None, None, None, C.byref(conflict)))
        k = n.value
        owners = np.zeros(k, dtype=np.int64)
        num, den, fac = (np.zeros(k * self.S) for _ in range(3))
        self._check(lib().bfvm_cell_sums(self._h, C.byref(n), _p(owners, i64p), _p(num), _p(den), _p(fac),
                                         C.byref(conflict)))
        r = (k, self.S)
        return owners, num.reshape(r), den.reshape(r), fac.reshape(r), bool(conflict.value)

    def device_ptr(self):
        p = C.c_void_p()
        self._check(lib().bfvm_field_device_ptr(self._h, C.byref(p)))
        return p.value

    def time_stages(self, steps):
        out = np.zeros(6)
        self._check(lib().bfvm_time_stages(self._h, int(steps), _p(out)))
        return dict(zip(("x", "y", "z", "dirichlet", "bulk", "cells"), out.tolist()))

    def time_sweep(self, axis, launches):
        """Average duration in microseconds of `launches` back-to-back sweeps
        along `axis` (measurement only: the field takes the extra sweeps)."""
        out = np.zeros(1)
        self._check(lib().bfvm_time_sweep(self._h, int(axis), int(launches), _p(out)))
        return float(out[0])

    def sweep_plan(self):
        """The kernel family of each axis' sweep and, when the x sweep runs
        in two halves, the tiles of the first halves of x and y."""
        k = np.zeros(3, dtype=np.int32)
        sp = np.zeros(4, dtype=np.int32)
        i32 = C.POINTER(C.c_int32)
        self._check(lib().bfvm_sweep_plan(self._h, _p(k, i32), _p(sp, i32)))
        return {"kernels": k.tolist(), "split_x": int(sp[0]), "split_y": int(sp[1]),
                "x_tiles": int(sp[2]), "y_tiles": int(sp[3])}

    def tuning(self):
        t = Tuning()
        self._check(lib().bfvm_get_tuning(self._h, C.byref(t)))
        return {"persistent_wgs_per_cu": t.persistent_wgs_per_cu,
                "force_lds_sweeps": t.force_lds_sweeps, "x_sweep": t.x_sweep,
                "pipeline": t.pipeline, "round_nearest": t.round_nearest,
                "lines_per_block": t.lines_per_block, "split_xy": t.split_xy}

    def set_tuning(self, persistent_wgs_per_cu=0, force_lds_sweeps=0, x_sweep=0, pipeline=0,
                   round_nearest=0, lines_per_block=0, split_xy=0):
        t = Tuning()
        t.persistent_wgs_per_cu = int(persistent_wgs_per_cu)
        t.force_lds_sweeps = int(force_lds_sweeps)
        t.x_sweep = int(x_sweep)
        t.pipeline = int(pipeline)
        t.round_nearest = int(round_nearest)
        t.lines_per_block = int(lines_per_block)
        t.split_xy = int(split_xy)
        self._check(lib().bfvm_set_tuning(self._h, C.byref(t)))


def _tuning_dict(t):
    return {"persistent_wgs_per_cu": t.persistent_wgs_per_cu, "force_lds_sweeps": t.force_lds_sweeps,
            "x_sweep": t.x_sweep, "pipeline": t.pipeline, "round_nearest": t.round_nearest,
            "lines_per_block": t.lines_per_block, "split_xy": t.split_xy}


class Batch:
    """Several Domains stepped together (a bfvm_batch).  The domains stay
    usable on their own.  Once one of them is closed the batch may only wait
    on its tickets and be closed."""

    def __init__(self, domains):
        self.domains = list(domains)
        arr = (C.c_void_p * len(self.domains))(*[d._h for d in self.domains])
        h = C.c_void_p()
        rc = lib().bfvm_batch_create(arr, len(self.domains), C.byref(h))
        if rc != OK:
            raise BfvmError(rc, lib().bfvm_batch_last_error(None).decode())
        self._h = h

    def _check(self, rc):
        if rc != OK:
            raise BfvmError(rc, lib().bfvm_batch_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            lib().bfvm_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, iterations=1, graph_steps=0):
        self._check(lib().bfvm_batch_step(self._h, int(iterations), int(graph_steps)))

    def sync(self):
        self._check(lib().bfvm_batch_sync(self._h))

    def mark(self):
        """Ticket for the work queued so far (wait() blocks until it has run)."""
        t = C.c_int64()
        self._check(lib().bfvm_batch_mark(self._h, C.byref(t)))
        return t.value

    def wait(self, ticket):
        self._check(lib().bfvm_batch_wait(self._h, int(ticket)))

    def query(self, ticket):
        """Whether ticket's work has run (does not block)."""
        d = C.c_int32()
        self._check(lib().bfvm_batch_query(self._h, int(ticket), C.byref(d)))
        return bool(d.value)

    def info(self):
        a, b = C.c_int32(), C.c_int32()
        self._check(lib().bfvm_batch_info(self._h, C.byref(a), C.byref(b)))
        return {"sweep_launches": a.value, "stage_launches": b.value}

    def shape_key(self):
        """What the tuning choice depends on: each domain's dimensions, grid,
        substrate count and agent count, in batch order."""
        return ";".join(f"{d.dims}:{'x'.join(map(str, d.grid))}:{d.S}:{d.n_agents}" for d in self.domains)

    def autotune(self, steps=5, cache=None):
        """Times the candidate tunings (bfvm_batch_autotune; the domains are
        left as they were) and applies the fastest to every domain.  cache: a
        JSON file of earlier choices keyed by shape_key(); a hit applies the
        stored choice without timing, a miss times and adds it (the file is
        replaced atomically)."""
        import json
        import os

        key = self.shape_key()
        table = {}
        if cache and os.path.exists(cache):
            with open(cache) as f:
                table = json.load(f)
            if key in table:
                t = dict(table[key])
                for d in self.domains:
                    own = d.tuning()
                    d.set_tuning(**{**t, "round_nearest": own["round_nearest"]})
                return t
        t = Tuning()
        self._check(lib().bfvm_batch_autotune(self._h, int(steps), C.byref(t)))
        choice = _tuning_dict(t)
        if cache:
            table[key] = {k: v for k, v in choice.items() if k != "round_nearest"}
            tmp = f"{cache}.{os.getpid()}.tmp"
            with open(tmp, "w") as f:
                json.dump(table, f, indent=1, sort_keys=True)
            os.replace(tmp, cache)
        return choice
