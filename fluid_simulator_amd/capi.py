"""ctypes binding of libfsim_hip.so (the C-ABI declared in include/fsim.h).

This is the Python view of the boundary a reference-side caller binds: the
structs mirror include/fsim.h field for field and every entry point is
prototyped.  Loading fails loudly when the HIP library has not been built;
there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfsim_hip.so")

# fsim_status
OK, ERR_INVALID, ERR_HIP, ERR_UNSUPPORTED, ERR_OOM, ERR_NUMERIC, ERR_TIMEOUT, ERR_NO_DEVICE = range(8)
# fsim_transfer (Simulator::P2G2PType, simulator.h:22-24)
PIC, FLIP, APIC = 0, 1, 2
# fsim_cell_type (MacGridCell::CellType, macGridCell.h:26-28)
WATER, AIR, SOLID = 0, 1, 2
# fsim_obstacle_kind (obstacles.hpp:7-65)
OBS_BOX, OBS_SPHERE, OBS_SOURCE, OBS_SINK = 0, 1, 2, 3
STOP_ABS_INF, STOP_REL_L2 = 0, 1
PRECOND_MIC0, PRECOND_LEARNED, SOLVER_BASIC_SOR = 0, 1, 2
# fsim_stage
(STAGE_SPAWN, STAGE_ADVECT, STAGE_PUSH_OUT, STAGE_STOP, STAGE_BIN, STAGE_P2G, STAGE_MARK, STAGE_OBSTACLES,
 STAGE_BORDERS, STAGE_POST_P2G, STAGE_SOLVE, STAGE_EXTRAPOLATE, STAGE_G2P, STAGE_PUSH_APART) = range(14)
# fsim_timer, named as the reference's stepDuration keys (simulator.cpp:31-41)
TIMER_KEYS = ("SimulateParticles", "PushParticlesApart", "PushParticlesOutOfObstacles", "P2GTransfer",
              "IncompressibilityPrep", "Incompressibility", "VelocityExtrapolation", "G2PTransfer")
# fsim_field
(FIELD_U, FIELD_V, FIELD_W, FIELD_U2, FIELD_V2, FIELD_W2, FIELD_LABEL, FIELD_DENSITY, FIELD_PRESSURE, FIELD_RHS,
 FIELD_PRECON, FIELD_CELL_COUNT, FIELD_CELL_START) = range(13)
FIELD_DTYPES = {FIELD_U: np.float32, FIELD_V: np.float32, FIELD_W: np.float32, FIELD_U2: np.float32,
                FIELD_V2: np.float32, FIELD_W2: np.float32, FIELD_LABEL: np.uint8, FIELD_DENSITY: np.float32,
                FIELD_PRESSURE: np.float64, FIELD_RHS: np.float32, FIELD_PRECON: np.float64,
                FIELD_CELL_COUNT: np.int32, FIELD_CELL_START: np.int32}
OP_APPLY_A, OP_APPLY_MIC0, OP_FACTOR_MIC0, OP_APPLY_MG = 0, 1, 2, 3


class GridDesc(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("cell_d", C.c_double),
                ("particle_r", C.c_double), ("two_d", C.c_int), ("cell_d_z", C.c_double)]


class Params(C.Structure):
    _fields_ = [("transfer", C.c_int), ("flip_ratio", C.c_float), ("gravity", C.c_float),
                ("gravity_enabled", C.c_int), ("push_apart_enabled", C.c_int), ("spawning_enabled", C.c_int),
                ("despawning_enabled", C.c_int), ("stop_particles", C.c_int), ("top_solid", C.c_int),
                ("pressure_k", C.c_double), ("average_pressure", C.c_double), ("max_iterations", C.c_int),
                ("pressure_enabled", C.c_int), ("fluid_density", C.c_double), ("residual_tolerance", C.c_double),
                ("stop_rule", C.c_int), ("precond", C.c_int), ("seed", C.c_uint64)]


class Obstacle(C.Structure):
    _fields_ = [("kind", C.c_int), ("pos", C.c_double * 3), ("speed", C.c_double * 3), ("size", C.c_double * 3),
                ("r", C.c_double), ("spawn_rate", C.c_double), ("spawn_speed", C.c_double),
                ("last_spawn_fraction", C.c_double)]


class GfxParticle(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("v", C.c_float), ("density", C.c_float)]


GFX_DTYPE = np.dtype([("pos", np.float32, 3), ("v", np.float32), ("density", np.float32)])


class StepStats(C.Structure):
    _fields_ = [("timer_ms", C.c_double * 8), ("step_ms", C.c_double), ("pcg_iterations", C.c_int),
                ("early_exit", C.c_int), ("fallback", C.c_int), ("reliable_updates", C.c_int),
                ("residual_abs_inf", C.c_double), ("residual_rel_l2", C.c_double), ("b_norm_l2", C.c_double),
                ("solve_ms", C.c_double), ("num_particles", C.c_int64), ("num_fluid_cells", C.c_int64),
                ("learned_iterations", C.c_int), ("learned_stop", C.c_int), ("solve_nan", C.c_int),
                ("learned_budget", C.c_int)]

    def as_dict(self) -> dict:
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "timer_ms"}
        d["timers_ms"] = {k: self.timer_ms[i] for i, k in enumerate(TIMER_KEYS)}
        return d


# Every entry point declared in include/fsim.h: name -> (restype, argtypes)
_P = C.c_void_p
PROTOTYPES = {
    "fsim_create": (C.c_int, [C.POINTER(GridDesc), C.POINTER(Params), C.POINTER(_P)]),
    "fsim_destroy": (C.c_int, [_P]),
    "fsim_last_error": (C.c_char_p, [_P]),
    "fsim_status_string": (C.c_char_p, [C.c_int]),
    "fsim_abi_version": (C.c_int, []),
    "fsim_set_params": (C.c_int, [_P, C.POINTER(Params)]),
    "fsim_get_params": (C.c_int, [_P, C.POINTER(Params)]),
    "fsim_set_particle_radius": (C.c_int, [_P, C.c_double]),
    "fsim_set_particles": (C.c_int, [_P, C.c_int64, _P, _P, _P]),
    "fsim_get_particles": (C.c_int, [_P, C.c_int64, _P, _P, _P, C.POINTER(C.c_int64)]),
    "fsim_num_particles": (C.c_int, [_P, C.POINTER(C.c_int64)]),
    "fsim_get_slot_ids": (C.c_int, [_P, C.c_int64, _P, C.POINTER(C.c_int64)]),
    "fsim_set_obstacles": (C.c_int, [_P, C.c_int, C.POINTER(Obstacle)]),
    "fsim_get_obstacles": (C.c_int, [_P, C.c_int, C.POINTER(Obstacle), C.POINTER(C.c_int)]),
    "fsim_step": (C.c_int, [_P, C.c_double, C.POINTER(StepStats)]),
    "fsim_run_stage": (C.c_int, [_P, C.c_int, C.c_double]),
    "fsim_get_stats": (C.c_int, [_P, C.POINTER(StepStats)]),
    "fsim_get_grid": (C.c_int, [_P, C.c_int, _P, C.c_size_t]),
    "fsim_set_grid": (C.c_int, [_P, C.c_int, _P, C.c_size_t]),
    "fsim_solver_apply": (C.c_int, [_P, C.c_int, C.c_double, _P, _P]),
    "fsim_learned_num_params": (C.c_int64, []),
    "fsim_learned_init": (C.c_int, [_P, C.c_uint64]),
    "fsim_learned_set_params": (C.c_int, [_P, _P]),
    "fsim_learned_get_params": (C.c_int, [_P, _P]),
    "fsim_learned_forward": (C.c_int, [_P, _P, _P, _P]),
    "fsim_learned_grad": (C.c_int, [_P, _P, _P, _P, _P, C.POINTER(C.c_double)]),
    "fsim_learned_train_batch": (C.c_int, [_P, C.c_int, _P, _P, _P, C.c_float, C.POINTER(C.c_double)]),
    "fsim_learned_forward_op": (C.c_int, [_P, _P, _P, _P]),
    "fsim_learned_grad_op": (C.c_int, [_P, _P, _P, _P, _P, C.POINTER(C.c_double)]),
    "fsim_learned_train_batch_op": (C.c_int, [_P, C.c_int, _P, _P, _P, C.c_float, C.POINTER(C.c_double)]),
    "fsim_mg_info": (C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "fsim_learned_harvest": (C.c_int, [_P, C.c_double, C.c_int, _P, _P, _P, C.POINTER(C.c_int)]),
    "fsim_scene_dam_break": (C.c_int, [C.POINTER(GridDesc), C.c_int, _P, C.c_int, C.POINTER(Obstacle), C.c_uint64,
                                       C.c_int64, _P, C.POINTER(C.c_int64)]),
    "fsim_gfx_snapshot": (C.c_int, [_P, C.c_int64, _P, C.POINTER(C.c_int64)]),
    "fsim_gfx_snapshot_noalloc": (C.c_int, [_P, C.c_int64, _P, C.POINTER(C.c_int64)]),
    "fsim_gfx_snapshot_noalloc_max": (C.c_int64, []),
    "fsim_set_option": (C.c_int, [_P, C.c_int, C.c_int]),
    "fsim_num_kernels": (C.c_int, []),
    "fsim_kernel_name": (C.c_char_p, [C.c_int]),
    "fsim_get_kernel_time": (C.c_int, [_P, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "fsim_reset_kernel_times": (C.c_int, [_P]),
}
(OPT_SWEEP_MODE, OPT_PROFILE, OPT_LEARNED_SMOOTHING, OPT_LEARNED_LOSS, OPT_LEARNED_CONV, OPT_LEARNED_REFRESH,
 OPT_LEARNED_MIC_COUNT, OPT_MG_COARSE, OPT_RELIABLE_DROP, OPT_MG_GRAPH, OPT_P2G_KERNEL,
 OPT_WAVE_SEND_DELAY) = range(12)
MG_COARSE_NETWORK, MG_COARSE_SMOOTH, MG_COARSE_ACCURATE = 0, 1, 2

_lib = None


def load(path: str = LIB_PATH) -> C.CDLL:
    """Load libfsim_hip.so (built by __graft_entry__.build()); raises if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build the HIP extension first (python -c "
                           "'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(path)
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class FsimError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"fsim status {status}: {msg}")
        self.status = status


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def make_params(**kw) -> Params:
    """Params with the reference defaults (simulator.h:30-39, macGrid.h:173-179),
    overridden by keyword arguments."""
    p = Params()
    defaults = dict(transfer=FLIP, flip_ratio=0.99, gravity=150.0, gravity_enabled=1, push_apart_enabled=0,
                    spawning_enabled=0, despawning_enabled=0, stop_particles=0, top_solid=0, pressure_k=2.0,
                    average_pressure=2.0, max_iterations=80, pressure_enabled=1, fluid_density=1.0,
                    residual_tolerance=1e-6, stop_rule=STOP_ABS_INF, precond=PRECOND_MIC0, seed=0)
    defaults.update(kw)
    for k, v in defaults.items():
        setattr(p, k, v)
    return p


def make_obstacle(kind: int, pos, speed=(0.0, 0.0, 0.0), size=(0.0, 0.0, 0.0), r=0.0, spawn_rate=0.0,
                  spawn_speed=0.0, last_spawn_fraction=0.0) -> Obstacle:
    o = Obstacle()
    o.kind = kind
    for i in range(3):
        o.pos[i] = pos[i]
        o.speed[i] = speed[i]
        o.size[i] = size[i]
    o.r, o.spawn_rate, o.spawn_speed, o.last_spawn_fraction = r, spawn_rate, spawn_speed, last_spawn_fraction
    return o


@dataclass
class Grid:
    """MacGrid's size and cell size; two_d: the reference's 2D mode (nz = 3, cell_d_z = the depth / 3)."""
    nx: int
    ny: int
    nz: int
    cell_d: float
    particle_r: float
    two_d: int = 0
    cell_d_z: float = 0.0

    def desc(self) -> GridDesc:
        return GridDesc(self.nx, self.ny, self.nz, self.cell_d, self.particle_r, self.two_d, self.cell_d_z)

    @property
    def cell_dz(self) -> float:
        """cellD.z (cell_d outside the 2D mode)."""
        return self.cell_d_z if self.two_d else self.cell_d

    @property
    def dims(self):
        return (self.nx * self.cell_d, self.ny * self.cell_d, self.nz * self.cell_dz)

    @property
    def nc(self) -> int:
        return self.nx * self.ny * self.nz

    @property
    def shape(self):
        return (self.nx, self.ny, self.nz)


def scene_dam_break(grid: Grid, blocks, obstacles=(), seed: int = 0) -> np.ndarray:
    """Dam-break particle positions (n, 3) float32 from the library's host generator."""
    lib = load()
    blk = np.ascontiguousarray(np.asarray(blocks, dtype=np.float64).reshape(-1, 6))
    obs = (Obstacle * max(len(obstacles), 1))(*obstacles)
    n = C.c_int64(0)
    d = grid.desc()
    rc = lib.fsim_scene_dam_break(C.byref(d), len(blk), _ptr(blk), len(obstacles), obs, seed, 0, None, C.byref(n))
    if rc:
        raise FsimError(rc, "scene generation failed")
    pos = np.empty((n.value, 3), np.float32)
    rc = lib.fsim_scene_dam_break(C.byref(d), len(blk), _ptr(blk), len(obstacles), obs, seed, n.value, _ptr(pos),
                                  C.byref(n))
    if rc:
        raise FsimError(rc, "scene generation failed")
    return pos


def learned_num_params() -> int:
    return int(load().fsim_learned_num_params())


class Simulation:
    """One fsim context: the device-resident state of one Simulator step."""

    def __init__(self, grid: Grid, params: Params):
        self.lib = load()
        self.grid = grid
        self._ctx = C.c_void_p()
        d = grid.desc()
        rc = self.lib.fsim_create(C.byref(d), C.byref(params), C.byref(self._ctx))
        if rc:
            raise FsimError(rc, self.lib.fsim_status_string(rc).decode())
        self.params = params

    def _check(self, rc: int):
        if rc:
            raise FsimError(rc, self.lib.fsim_last_error(self._ctx).decode())

    def close(self):
        if self._ctx:
            self.lib.fsim_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params: Params):
        self._check(self.lib.fsim_set_params(self._ctx, C.byref(params)))
        self.params = params

    def set_particle_radius(self, r: float):
        self._check(self.lib.fsim_set_particle_radius(self._ctx, r))

    def set_particles(self, pos, vel=None, c=None):
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
        vel = None if vel is None else np.ascontiguousarray(vel, dtype=np.float32).reshape(-1, 3)
        c = None if c is None else np.ascontiguousarray(c, dtype=np.float32).reshape(-1, 9)
        self._check(self.lib.fsim_set_particles(self._ctx, len(pos), _ptr(pos), _ptr(vel), _ptr(c)))

    def num_particles(self) -> int:
        n = C.c_int64()
        self._check(self.lib.fsim_num_particles(self._ctx, C.byref(n)))
        return n.value

    def slot_ids(self) -> np.ndarray:
        n = self.num_particles()
        ids = np.empty(max(n, 1), np.int32)
        got = C.c_int64()
        self._check(self.lib.fsim_get_slot_ids(self._ctx, n, _ptr(ids), C.byref(got)))
        return ids[:n]

    def get_particles(self):
        n = self.num_particles()
        pos = np.empty((n, 3), np.float32)
        vel = np.empty((n, 3), np.float32)
        c = np.empty((n, 9), np.float32)
        got = C.c_int64()
        self._check(self.lib.fsim_get_particles(self._ctx, n, _ptr(pos), _ptr(vel), _ptr(c), C.byref(got)))
        return pos, vel, c

    def set_obstacles(self, obstacles):
        arr = (Obstacle * max(len(obstacles), 1))(*obstacles)
        self._check(self.lib.fsim_set_obstacles(self._ctx, len(obstacles), arr))

    def get_obstacles(self):
        n = C.c_int()
        self._check(self.lib.fsim_get_obstacles(self._ctx, 0, None, C.byref(n)))
        arr = (Obstacle * max(n.value, 1))()
        self._check(self.lib.fsim_get_obstacles(self._ctx, n.value, arr, C.byref(n)))
        return list(arr)[: n.value]

    def step(self, dt: float) -> StepStats:
        st = StepStats()
        self._check(self.lib.fsim_step(self._ctx, dt, C.byref(st)))
        return st

    def run_stage(self, stage: int, dt: float = 0.0):
        self._check(self.lib.fsim_run_stage(self._ctx, stage, dt))

    def stats(self) -> StepStats:
        st = StepStats()
        self._check(self.lib.fsim_get_stats(self._ctx, C.byref(st)))
        return st

    def get_grid(self, field: int) -> np.ndarray:
        n = self.grid.nc + (1 if field == FIELD_CELL_START else 0)
        out = np.empty(n, FIELD_DTYPES[field])
        self._check(self.lib.fsim_get_grid(self._ctx, field, _ptr(out), out.nbytes))
        return out

    def set_grid(self, field: int, values: np.ndarray):
        a = np.ascontiguousarray(values, dtype=FIELD_DTYPES[field]).reshape(-1)
        self._check(self.lib.fsim_set_grid(self._ctx, field, _ptr(a), a.nbytes))

    def gfx_snapshot(self, noalloc: bool = False) -> np.ndarray:
        """SimulationManager's ParticleGfxData of every particle, in particle order (GFX_DTYPE records).
        noalloc: through fsim_gfx_snapshot_noalloc (fixed per-GPU staging, no device allocation)."""
        n = self.num_particles()
        out = np.zeros(max(n, 1), GFX_DTYPE)
        got = C.c_int64()
        fn = self.lib.fsim_gfx_snapshot_noalloc if noalloc else self.lib.fsim_gfx_snapshot
        self._check(fn(self._ctx, n, _ptr(out), C.byref(got)))
        return out[:n]

    def set_option(self, option: int, value: int):
        self._check(self.lib.fsim_set_option(self._ctx, option, value))

    def kernel_times(self) -> dict:
        """{kernel name: (total ms, launches)} accumulated while profiling is on."""
        out = {}
        for k in range(self.lib.fsim_num_kernels()):
            ms, n = C.c_double(), C.c_int64()
            self._check(self.lib.fsim_get_kernel_time(self._ctx, k, C.byref(ms), C.byref(n)))
            out[self.lib.fsim_kernel_name(k).decode()] = (ms.value, n.value)
        return out

    def reset_kernel_times(self):
        self._check(self.lib.fsim_reset_kernel_times(self._ctx))

    def solver_apply(self, op: int, dt: float, vec: np.ndarray | None = None) -> np.ndarray:
        out = np.empty(self.grid.nc, np.float64)
        v = None if vec is None else np.ascontiguousarray(vec, dtype=np.float64).reshape(-1)
        self._check(self.lib.fsim_solver_apply(self._ctx, op, dt, _ptr(v), _ptr(out)))
        return out

    # ---- learned preconditioner ----
    def learned_init(self, seed: int = 0):
        self._check(self.lib.fsim_learned_init(self._ctx, seed))

    def learned_set_params(self, params: np.ndarray):
        p = np.ascontiguousarray(params, dtype=np.float32).reshape(-1)
        if p.size != learned_num_params():
            raise ValueError(f"expected {learned_num_params()} parameters, got {p.size}")
        self._check(self.lib.fsim_learned_set_params(self._ctx, _ptr(p)))

    def learned_get_params(self) -> np.ndarray:
        p = np.empty(learned_num_params(), np.float32)
        self._check(self.lib.fsim_learned_get_params(self._ctx, _ptr(p)))
        return p

    def _sample(self, b, label):
        b = np.ascontiguousarray(b, dtype=np.float32).reshape(-1)
        lab = np.ascontiguousarray(label, dtype=np.uint8).reshape(-1)
        if b.size != self.grid.nc or lab.size != self.grid.nc:
            raise ValueError("b and label need one value per cell")
        return b, lab

    def learned_forward(self, b: np.ndarray, label: np.ndarray) -> np.ndarray:
        b, lab = self._sample(b, label)
        out = np.empty(self.grid.nc, np.float32)
        self._check(self.lib.fsim_learned_forward(self._ctx, _ptr(b), _ptr(lab), _ptr(out)))
        return out

    def learned_grad(self, b: np.ndarray, label: np.ndarray, p: np.ndarray):
        b, lab = self._sample(b, label)
        pv = np.ascontiguousarray(p, dtype=np.float64).reshape(-1)
        g = np.empty(learned_num_params(), np.float32)
        loss = C.c_double(0.0)
        self._check(self.lib.fsim_learned_grad(self._ctx, _ptr(b), _ptr(lab), _ptr(pv), _ptr(g), C.byref(loss)))
        return loss.value, g

    def _op(self, op):
        o = np.ascontiguousarray(op, dtype=np.float32).reshape(-1)
        if o.size % (4 * self.grid.nc):
            raise ValueError("an operator has 4 values per cell: diag, cx, cy, cz")
        return o

    def learned_forward_op(self, b: np.ndarray, op: np.ndarray) -> np.ndarray:
        bb = np.ascontiguousarray(b, dtype=np.float32).reshape(-1)
        out = np.empty(self.grid.nc, np.float32)
        self._check(self.lib.fsim_learned_forward_op(self._ctx, _ptr(bb), _ptr(self._op(op)), _ptr(out)))
        return out

    def learned_grad_op(self, b: np.ndarray, op: np.ndarray, p: np.ndarray):
        bb = np.ascontiguousarray(b, dtype=np.float32).reshape(-1)
        pv = np.ascontiguousarray(p, dtype=np.float64).reshape(-1)
        g = np.empty(learned_num_params(), np.float32)
        loss = C.c_double(0.0)
        self._check(self.lib.fsim_learned_grad_op(self._ctx, _ptr(bb), _ptr(self._op(op)), _ptr(pv), _ptr(g),
                                                  C.byref(loss)))
        return loss.value, g

    def learned_train_batch_op(self, b: np.ndarray, op: np.ndarray, p: np.ndarray, lr: float) -> float:
        n = len(b)
        bb = np.ascontiguousarray(b, dtype=np.float32).reshape(n, -1)
        pv = np.ascontiguousarray(p, dtype=np.float64).reshape(n, -1)
        o = self._op(op)
        if o.size != 4 * bb.size:
            raise ValueError("one operator per sample")
        loss = C.c_double(0.0)
        self._check(self.lib.fsim_learned_train_batch_op(self._ctx, n, _ptr(bb), _ptr(o), _ptr(pv), C.c_float(lr),
                                                         C.byref(loss)))
        return loss.value

    def mg_info(self):
        """(number of multigrid levels, coarsest level's dims)"""
        lv = C.c_int()
        d = (C.c_int * 3)()
        self._check(self.lib.fsim_mg_info(self._ctx, C.byref(lv), d))
        return lv.value, tuple(d)

    def learned_harvest(self, dt: float, capacity: int):
        """Coarsest-level problems of one learned-multigrid solve of the current state (after the
        preparation stages): (b[n, Nc'], op[4, Nc'], x[n, Nc']) on the coarse level's grid."""
        _, dims = self.mg_info()
        nc = int(np.prod(dims))
        b = np.empty((max(capacity, 1), nc), np.float32)
        x = np.empty((max(capacity, 1), nc), np.float32)
        op = np.empty((4, nc), np.float32)
        n = C.c_int()
        self._check(self.lib.fsim_learned_harvest(self._ctx, dt, capacity, _ptr(b), _ptr(op), _ptr(x), C.byref(n)))
        return b[: n.value], op, x[: n.value], dims

    def learned_train_batch(self, b: np.ndarray, label: np.ndarray, p: np.ndarray, lr: float) -> float:
        n = len(b)
        bb = np.ascontiguousarray(b, dtype=np.float32).reshape(n, -1)
        lab = np.ascontiguousarray(label, dtype=np.uint8).reshape(n, -1)
        pv = np.ascontiguousarray(p, dtype=np.float64).reshape(n, -1)
        loss = C.c_double(0.0)
        self._check(self.lib.fsim_learned_train_batch(self._ctx, n, _ptr(bb), _ptr(lab), _ptr(pv), C.c_float(lr),
                                                      C.byref(loss)))
        return loss.value
