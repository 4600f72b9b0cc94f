"""Shared helpers for the parity tests (GPU path vs the FP64 oracle)."""
from __future__ import annotations

import numpy as np

from fluid_simulator_amd import capi, scenes


def field_rel_err(gpu: np.ndarray, ref: np.ndarray) -> float:
    """max |gpu - ref| / max |ref|: error relative to the field's scale."""
    ref = np.asarray(ref, np.float64)
    scale = max(np.abs(ref).max(), 1e-30)
    return float(np.abs(np.asarray(gpu, np.float64) - ref).max() / scale)


def ulp_mismatch(gpu32: np.ndarray, ref64: np.ndarray, ulps: float = 1.0) -> int:
    """Number of entries where the FP32 value is further than `ulps` FP32
    ulps from the FP64 reference value."""
    ref64 = np.asarray(ref64, np.float64)
    tol = ulps * np.spacing(np.abs(ref64).astype(np.float32)).astype(np.float64)
    return int(np.count_nonzero(np.abs(gpu32.astype(np.float64) - ref64) > tol + 1e-300))


def random_particles(grid: capi.Grid, n: int, seed: int, vmax: float = 1.0, apic: bool = False):
    """FP32-representable random particles strictly inside the clamp box."""
    rng = np.random.default_rng(seed)
    lo = grid.cell_d + 1.01 * grid.particle_r
    dims = np.array(grid.shape) * grid.cell_d
    pos = rng.uniform(lo, dims - lo, size=(n, 3)).astype(np.float32)
    vel = rng.uniform(-vmax, vmax, size=(n, 3)).astype(np.float32)
    c = rng.uniform(-vmax, vmax, size=(n, 9)).astype(np.float32) if apic else None
    return pos, vel, c


def pair(grid: capi.Grid, prm: capi.Params, pos, vel=None, c=None, obstacles=()):
    """A GPU simulation and an oracle holding the same FP32-representable state."""
    import oracle

    sim = capi.Simulation(grid, prm)
    ref = oracle.Oracle(grid, prm)
    sim.set_particles(pos, vel, c)
    ref.set_particles(np.asarray(pos, np.float64), None if vel is None else np.asarray(vel, np.float64),
                      None if c is None else np.asarray(c, np.float64))
    if obstacles:
        sim.set_obstacles(list(obstacles))
        ref.set_obstacles(list(obstacles))
    return sim, ref


def sync_oracle_from_gpu(sim: capi.Simulation, ref) -> None:
    pos, vel, c = sim.get_particles()
    ref.set_particles(pos.astype(np.float64), vel.astype(np.float64), c.astype(np.float64))
    ref.set_obstacles(sim.get_obstacles())


def dam_state(n: int = 16, transfer=capi.FLIP, steps: int = 2, sphere: bool = False, **kw):
    """A dam-break state a few steps in (so velocities are non-trivial), as
    FP32 particle arrays and the params used."""
    sc = scenes.dam_break(n, transfer, sphere=sphere)
    prm = scenes.params(transfer=transfer, **kw)
    sim = capi.Simulation(sc.grid, prm)
    sim.set_particles(sc.particles())
    if sc.obstacles:
        sim.set_obstacles(sc.obstacles)
    for _ in range(steps):
        sim.step(scenes.DT)
    pos, vel, c = sim.get_particles()
    sim.close()
    return sc, prm, pos, vel, c
