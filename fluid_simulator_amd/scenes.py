"""Synthetic dam-break scenes of the benchmark configurations.

Common choices (the reference leaves them open): domain [0,1]^3 built as
MacGrid(targetDimensions=(1,1,1), resolution=N), particle radius dx/4,
push-apart off (the reference's 3D preset, simulationGui.cpp:57), pressure
(density) term off, open top, rho = 1, gravity -9.81 on y, fixed dt = 1/240,
PCG cap raised so that the 1e-6 tolerance is reached.
"""
from __future__ import annotations

from dataclasses import dataclass, field

from . import capi

DT = 1.0 / 240.0


def grid(n: int) -> capi.Grid:
    return capi.Grid(n, n, n, 1.0 / n, 0.25 / n)


DEPTH_2D = 0.1


def grid_2d(n: int, depth: float = DEPTH_2D) -> capi.Grid:
    """The 2D mode: MacGrid((1, 1, depth), resolution n, twoD) is n x n x 3
    cells with cellD.z = depth / 3 (macGrid.cpp:9-13); particles stay at
    z = depth / 2."""
    return capi.Grid(n, n, 3, 1.0 / n, 0.25 / n, two_d=1, cell_d_z=depth / 3)


def params(transfer=capi.FLIP, stop_rule=capi.STOP_REL_L2, tol=1e-6, max_iterations=10000, **kw) -> capi.Params:
    base = dict(transfer=transfer, flip_ratio=0.95, gravity=-9.81, gravity_enabled=1, push_apart_enabled=0,
                pressure_enabled=0, top_solid=0, fluid_density=1.0, residual_tolerance=tol,
                max_iterations=max_iterations, stop_rule=stop_rule)
    base.update(kw)
    return capi.make_params(**base)


@dataclass
class Scene:
    name: str
    n: int
    transfer: int
    blocks: list
    obstacles: list = field(default_factory=list)
    seed: int = 0
    steps: int = 100
    two_d: bool = False

    @property
    def grid(self) -> capi.Grid:
        return grid_2d(self.n) if self.two_d else grid(self.n)

    def particles(self):
        return capi.scene_dam_break(self.grid, self.blocks, self.obstacles, self.seed)


def dam_break(n: int, transfer=capi.FLIP, seed=0, sphere=False, steps=100) -> Scene:
    """Block x < 0.4, y < 0.6 (full z), 8 particles per fluid cell."""
    obs = [capi.make_obstacle(capi.OBS_SPHERE, (0.5, 0.5, 0.5), r=0.1)] if sphere else []
    return Scene(f"dam_break_{n}", n, transfer, [[0.0, 0.0, 0.0, 0.4, 0.6, 1.0]], obs, seed, steps)


def dam_break_2d(n: int, transfer=capi.FLIP, seed=0, sphere=False, steps=100) -> Scene:
    """The 2D mode's dam break: block x < 0.4, y < 0.6 of the n x n x 3 grid,
    4 particles per fluid cell at z = depth / 2; the 2D preset turns push-apart
    on (simulationGui.cpp:40)."""
    obs = [capi.make_obstacle(capi.OBS_SPHERE, (0.5, 0.5, DEPTH_2D / 2), r=0.1)] if sphere else []
    return Scene(f"dam_break_2d_{n}", n, transfer, [[0.0, 0.0, 0.0, 0.4, 0.6, DEPTH_2D]], obs, seed, steps, True)


def double_dam_break(n: int, transfer=capi.FLIP, seed=1, steps=100) -> Scene:
    """Two opposing blocks x < 0.4 and x > 0.6, y < 0.6 (full z)."""
    return Scene(f"double_dam_break_{n}", n, transfer,
                 [[0.0, 0.0, 0.0, 0.4, 0.6, 1.0], [0.6, 0.0, 0.0, 1.0, 0.6, 1.0]], [], seed, steps)


def boxes_and_source(n: int, transfer=capi.FLIP, seed=5, steps=500, per_step: int = 2000) -> Scene:
    """The double dam break with 3-8 solid boxes and a particle source.

    Box count, edge lengths (each U[0.05, 0.2]) and centres come from a seeded
    numpy generator: centre x, z ~ U[0.2, 0.8], y ~ U[size_y / 2 + 0.05, 0.45]
    (in the water's path).  The source (SphericalParticleSource,
    obstacles.hpp:45-56) sits at (0.5, 0.8, 0.5) with radius 0.04 and spawns
    per_step particles per step at DT (rate per_step / DT, simulator.cpp:
    112-114), launched outwards at 0.5 m/s.  Cells inside a box or the source
    start without particles."""
    import numpy as np

    rng = np.random.default_rng(seed)
    nbox = int(rng.integers(3, 9))
    obs = []
    for _ in range(nbox):
        size = rng.uniform(0.05, 0.2, 3)
        ctr = (rng.uniform(0.2, 0.8), rng.uniform(size[1] / 2 + 0.05, 0.45), rng.uniform(0.2, 0.8))
        obs.append(capi.make_obstacle(capi.OBS_BOX, ctr, size=tuple(float(v) for v in size)))
    obs.append(capi.make_obstacle(capi.OBS_SOURCE, (0.5, 0.8, 0.5), r=0.04, spawn_rate=per_step / DT,
                                  spawn_speed=0.5))
    return Scene(f"boxes_and_source_{n}", n, transfer,
                 [[0.0, 0.0, 0.0, 0.4, 0.6, 1.0], [0.6, 0.0, 0.0, 1.0, 0.6, 1.0]], obs, seed, steps)


# The benchmark configurations: 32^3 CPU-path case, 128^3 APIC with a sphere,
# 256^3 FLIP double dam break, and the 256^3 learned run with boxes and a
# particle source.
CONFIG_32_FLIP = lambda: dam_break(32, capi.FLIP, seed=0, steps=100)  # noqa: E731
CONFIG_128_APIC_SPHERE = lambda: dam_break(128, capi.APIC, seed=0, sphere=True, steps=200)  # noqa: E731
CONFIG_256_FLIP_DOUBLE = lambda: double_dam_break(256, capi.FLIP, seed=1, steps=100)  # noqa: E731
CONFIG_256_BOXES_SOURCE = lambda: boxes_and_source(256, capi.FLIP, seed=5, steps=500)  # noqa: E731
