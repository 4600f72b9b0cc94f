"""ctypes binding of the FP64 CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may
import this package, and only as the checker or the CPU baseline.  The
product (fluid_simulator_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from fluid_simulator_amd.capi import GridDesc, Params, Obstacle, StepStats, Grid

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfsim_oracle.so")

_P = C.c_void_p
_PROTOS = {
    "fo_create": (_P, [C.POINTER(GridDesc), C.POINTER(Params)]),
    "fo_destroy": (None, [_P]),
    "fo_set_threads": (None, [_P, C.c_int]),
    "fo_set_params": (None, [_P, C.POINTER(Params)]),
    "fo_set_particle_radius": (None, [_P, C.c_double]),
    "fo_set_particles": (C.c_int, [_P, C.c_int64, _P, _P, _P]),
    "fo_num_particles": (C.c_int64, [_P]),
    "fo_get_particles": (None, [_P, _P, _P, _P]),
    "fo_set_obstacles": (None, [_P, C.c_int, C.POINTER(Obstacle)]),
    "fo_get_obstacles": (C.c_int, [_P, C.POINTER(Obstacle)]),
    "fo_step": (None, [_P, C.c_double, C.POINTER(StepStats)]),
    "fo_run_stage": (C.c_int, [_P, C.c_int, C.c_double]),
    "fo_get_stats": (None, [_P, C.POINTER(StepStats)]),
    "fo_get_grid": (C.c_int, [_P, C.c_int, _P]),
    "fo_set_grid": (C.c_int, [_P, C.c_int, _P]),
    "fo_solver_apply": (C.c_int, [_P, C.c_int, C.c_double, _P, _P]),
    "fo_bin": (None, [_P, _P, _P, _P]),
    "fo_splitmix64": (C.c_uint64, [C.c_uint64, C.c_uint64]),
    "fo_gfx_snapshot": (None, [_P, _P]),
}

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        lib = C.CDLL(LIB_PATH)
        for k, (res, args) in _PROTOS.items():
            f = getattr(lib, k)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


# field ids / dtypes of the oracle's get/set (float fields are FP64 here)
_DT = {6: np.uint8}


class Oracle:
    """FP64 CPU restatement of the reference's Simulator step."""

    def __init__(self, grid: Grid, params: Params, threads: int = 1):
        self.lib = load()
        self.grid = grid
        d = grid.desc()
        self._s = self.lib.fo_create(C.byref(d), C.byref(params))
        if not self._s:
            raise RuntimeError("oracle creation failed")
        self.lib.fo_set_threads(self._s, threads)

    def close(self):
        if self._s:
            self.lib.fo_destroy(self._s)
            self._s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_threads(self, n: int):
        self.lib.fo_set_threads(self._s, n)

    def set_params(self, p: Params):
        self.lib.fo_set_params(self._s, C.byref(p))

    def set_particle_radius(self, r: float):
        self.lib.fo_set_particle_radius(self._s, r)

    def set_particles(self, pos, vel=None, c=None):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        vel = None if vel is None else np.ascontiguousarray(vel, dtype=np.float64).reshape(-1, 3)
        c = None if c is None else np.ascontiguousarray(c, dtype=np.float64).reshape(-1, 9)
        if self.lib.fo_set_particles(self._s, len(pos), _ptr(pos), _ptr(vel), _ptr(c)):
            raise MemoryError("oracle particle allocation failed")

    def num_particles(self) -> int:
        return self.lib.fo_num_particles(self._s)

    def get_particles(self):
        n = self.num_particles()
        pos = np.empty((n, 3))
        vel = np.empty((n, 3))
        c = np.empty((n, 9))
        self.lib.fo_get_particles(self._s, _ptr(pos), _ptr(vel), _ptr(c))
        return pos, vel, c

    def set_obstacles(self, obstacles):
        arr = (Obstacle * max(len(obstacles), 1))(*obstacles)
        self.lib.fo_set_obstacles(self._s, len(obstacles), arr)

    def get_obstacles(self):
        arr = (Obstacle * 64)()
        n = self.lib.fo_get_obstacles(self._s, arr)
        return list(arr)[:n]

    def step(self, dt: float) -> StepStats:
        st = StepStats()
        self.lib.fo_step(self._s, dt, C.byref(st))
        return st

    def run_stage(self, stage: int, dt: float = 0.0):
        if self.lib.fo_run_stage(self._s, stage, dt):
            raise ValueError(f"unknown stage {stage}")

    def stats(self) -> StepStats:
        st = StepStats()
        self.lib.fo_get_stats(self._s, C.byref(st))
        return st

    def get_grid(self, field: int) -> np.ndarray:
        out = np.empty(self.grid.nc, _DT.get(field, np.float64))
        if self.lib.fo_get_grid(self._s, field, _ptr(out)):
            raise ValueError(f"field {field} not available in the oracle")
        return out

    def set_grid(self, field: int, values):
        a = np.ascontiguousarray(values, dtype=_DT.get(field, np.float64)).reshape(-1)
        if self.lib.fo_set_grid(self._s, field, _ptr(a)):
            raise ValueError(f"field {field} not settable in the oracle")

    def solver_apply(self, op: int, dt: float, vec=None) -> np.ndarray:
        out = np.empty(self.grid.nc)
        v = None if vec is None else np.ascontiguousarray(vec, dtype=np.float64).reshape(-1)
        if self.lib.fo_solver_apply(self._s, op, dt, _ptr(v), _ptr(out)):
            raise ValueError("unknown solver op")
        return out

    def gfx_snapshot(self) -> np.ndarray:
        """(n, 5): pos, |v|, density of each particle (simulationManager.cpp:223-231)."""
        n = self.num_particles()
        out = np.empty((max(n, 1), 5))
        self.lib.fo_gfx_snapshot(self._s, _ptr(out))
        return out[:n]

    def bin(self):
        nc = self.grid.nc
        n = self.num_particles()
        count = np.empty(nc, np.int32)
        start = np.empty(nc + 1, np.int32)
        perm = np.empty(max(n, 1), np.int64)
        self.lib.fo_bin(self._s, _ptr(count), _ptr(start), _ptr(perm))
        return count, start, perm[:n]


def splitmix64(seed: int, n: int) -> int:
    return load().fo_splitmix64(seed, n)
