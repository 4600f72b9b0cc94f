"""fsim_gfx_snapshot_noalloc: the renderer snapshot through the fixed
per-GPU scratch array, checked against fsim_gfx_snapshot (the same kernel
with the context's own staging buffer) on two calls in a row, the second
with fewer particles, and the size limit."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from fluid_simulator_amd import capi, scenes

pytestmark = pytest.mark.gpu


def say(msg: str):
    print(msg, flush=True)
    sys.stdout.flush()


def test_gfx_snapshot_noalloc_matches_the_allocating_entry_point(gpu):
    sc = scenes.dam_break(32, capi.FLIP)
    say("case: snapshot without allocation, 32^3 dam break after 2 steps, then a third of the particles")
    sim = capi.Simulation(sc.grid, scenes.params())
    pos = sc.particles()
    sim.set_particles(pos)
    for _ in range(2):
        sim.step(scenes.DT)
    a = sim.gfx_snapshot()
    b = sim.gfx_snapshot(noalloc=True)
    assert len(a) == len(b) == len(pos)
    assert a.tobytes() == b.tobytes()
    assert np.all(a["density"] > 0) and np.any(a["v"] > 0)
    # second call: fewer particles on the same GPU (the scratch still holds the longer first snapshot)
    p2, v2, _ = sim.get_particles()
    m = len(p2) // 3
    sim.set_particles(p2[:m], v2[:m])
    sim.run_stage(capi.STAGE_P2G, scenes.DT)
    a2 = sim.gfx_snapshot()
    b2 = sim.gfx_snapshot(noalloc=True)
    assert len(a2) == len(b2) == m
    assert a2.tobytes() == b2.tobytes()
    assert np.array_equal(b2["pos"], p2[:m])
    sim.close()


def test_gfx_snapshot_noalloc_rejects_more_than_its_maximum(gpu):
    cap = int(capi.load().fsim_gfx_snapshot_noalloc_max())
    assert cap == (64 << 20) // capi.GFX_DTYPE.itemsize
    grid = scenes.grid(16)
    say(f"case: snapshot without allocation over its maximum ({cap} + 1 particles)")
    rng = np.random.default_rng(3)
    lo = grid.cell_d + 1.01 * grid.particle_r
    pos = rng.uniform(lo, 1.0 - lo, (cap + 1, 3)).astype(np.float32)
    sim = capi.Simulation(grid, scenes.params())
    sim.set_particles(pos)
    with pytest.raises(capi.FsimError, match="more than") as e:
        sim.gfx_snapshot(noalloc=True)
    assert e.value.status == capi.ERR_INVALID
    assert len(sim.gfx_snapshot()) == cap + 1   # the allocating entry point takes any count
    sim.close()
