"""The benchmark configurations under the per-step parity bar: the GPU path
against the FP64 oracle, stage by stage (tests/helpers.py staged_step).

Every stage starts both backends from the same inputs (the oracle takes the
GPU's FP32 particles or grid fields after each stage), so a stage's error is
its own and the bar is the per-step one: labels, binning counts, offsets and
the permutation bit-exact; grid velocities (P2G, boundary conditions,
projection, extrapolation), density, right-hand side and pressure within
1e-5 of the field's scale; G2P velocities within 1e-5.  The solves of these
tests stop at a 1e-10 relative residual (1e-9 at 256^3): two solves stopped
at the configurations' 1e-6 differ by up to cond(A) 1e-6 whatever their
arithmetic, which would measure the stop rule instead of the
implementations.  Below about 1e-7 the FP32 iterations lean on the FP64
reliable updates, which add iterations, so at 1e-10 the GPU's count is held
to [oracle - 2, 1.5 oracle + 5].  The iteration counts are compared at the
configurations' own 1e-6 by separate runs of the same staged steps (the same
right-hand side on both sides), with the exact checks.  FP32 vectors (r, z,
s, q) cost iterations against the FP64 oracle: at the first step, +1 to +3 at
32^3, +2 at 256^3 (145 vs 143) and +9 at 128^3 (74 vs 65).  The reliable-
update threshold is at its best there: true residuals after every 10^-3
drop; 10^-2, 10^-4 and 10^-5 give 75-78 and 161-173 iterations, and none
before the tolerance 87 and 193.  The bar is max(3, 20 %)."""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

from fluid_simulator_amd import capi, scenes
from helpers import STAGE_EXACT_KEYS, STAGE_FIELD_KEYS, staged_step

pytestmark = pytest.mark.gpu
DT = scenes.DT
BAR = 1e-5


def say(msg: str):
    print(msg, flush=True)
    sys.stdout.flush()


def check_stages(e: dict, step: int, bar: float = BAR, counts_only: bool = False):
    for k in STAGE_EXACT_KEYS:
        assert e.get(k, 0) == 0, (step, k, e[k])
    g, o = e["pcg_gpu"], e["pcg_oracle"]
    if counts_only:   # a 1e-6 solve: iteration counts within max(3, 20 %)
        assert abs(g - o) <= max(3, 0.2 * o), (step, g, o)
        return
    for k in STAGE_FIELD_KEYS:
        assert e[k] < bar, (step, k, e[k])
    for k in ("advect_pos_abs", "push_apart_pos_abs", "push_out_pos_abs", "g2p_c"):
        assert e.get(k, 0.0) < bar, (step, k, e)
    assert o - 2 <= g <= 1.5 * o + 5, (step, g, o)   # a 1e-10 solve


def worst(errs: list) -> str:
    keys = STAGE_FIELD_KEYS + ("advect_pos_abs", "push_apart_pos_abs", "push_out_pos_abs", "g2p_c")
    return ", ".join(f"{k} {max(e.get(k, 0.0) for e in errs):.1e}" for k in keys)


def run_staged(sc, prm, steps: int, threads: int, start_steps: int = 0, spawn: bool = False,
               counts_only: bool = False):
    import oracle

    sim = capi.Simulation(sc.grid, prm)
    if sc.obstacles:
        sim.set_obstacles(sc.obstacles)
    sim.set_particles(sc.particles())
    for _ in range(start_steps):
        sim.step(DT)
    ref = oracle.Oracle(sc.grid, prm, threads=threads)
    errs = []
    for k in range(steps):
        e = staged_step(sim, ref, DT, spawn=spawn, push_apart=bool(prm.push_apart_enabled))
        errs.append(e)
        if k % 10 == 0 or k == steps - 1:
            say(f"  step {k}: pcg {e['pcg_gpu']} / {e['pcg_oracle']} (oracle), p {e['pressure']:.1e}, "
                f"p2g {e['p2g']:.1e}, g2p {e['g2p_vel']:.1e}")
        check_stages(e, k, counts_only=counts_only)
    if counts_only:
        d = [e["pcg_gpu"] - e["pcg_oracle"] for e in errs]
        say(f"  iteration count GPU - oracle over {steps} steps: min {min(d)}, max {max(d)}")
    else:
        say(f"  worst over {steps} steps: {worst(errs)}")
    ref.close()
    sim.close()
    return errs


def test_config_32_flip_dam_break_100_steps_staged(gpu):
    """32^3 FLIP dam break (x < 0.4, y < 0.6, seed 0, 51,840 particles), all 100 steps."""
    say("case: 32^3 FLIP dam break, 100 steps, stage by stage against the oracle")
    sc = scenes.CONFIG_32_FLIP()
    run_staged(sc, scenes.params(transfer=sc.transfer, tol=1e-10), sc.steps, threads=min(16, os.cpu_count() or 1))


@pytest.mark.timeout(900)
def test_config_128_apic_sphere_20_steps_staged(gpu):
    """128^3 APIC dam break with the static sphere (3.83M particles), 20 consecutive steps from the
    start, stage by stage."""
    say("case: 128^3 APIC dam break with a sphere, 20 steps, stage by stage against the oracle")
    sc = scenes.CONFIG_128_APIC_SPHERE()
    run_staged(sc, scenes.params(transfer=sc.transfer, tol=1e-10), 20, threads=min(16, os.cpu_count() or 1))


@pytest.mark.timeout(900)
def test_config_256_flip_double_dam_break_one_step_staged(gpu):
    """256^3 FLIP double dam break (62.8M particles), the first step stage by stage: binning and labels
    bit-exact at the full size, pressure and faces at 1e-5."""
    say("case: 256^3 FLIP double dam break, first step, stage by stage against the oracle")
    sc = scenes.CONFIG_256_FLIP_DOUBLE()
    run_staged(sc, scenes.params(transfer=sc.transfer, tol=1e-9), 1, threads=min(16, os.cpu_count() or 1))


@pytest.mark.timeout(900)
@pytest.mark.parametrize("config", ["32", "128", "256"])
def test_config_iteration_counts_at_1e_6(gpu, config):
    """The configurations' own tolerance, 1e-6: the same staged steps (identical right-hand sides on both
    sides), exact checks and PCG iteration counts within max(3, 20 %).  32^3: all 100 steps; 128^3: the
    first 5 steps; 256^3: the first step."""
    sc, steps = {"32": (scenes.CONFIG_32_FLIP, 100), "128": (scenes.CONFIG_128_APIC_SPHERE, 5),
                 "256": (scenes.CONFIG_256_FLIP_DOUBLE, 1)}[config]
    sc = sc()
    say(f"case: {sc.name}, {steps} staged steps at 1e-6, iteration counts against the oracle")
    run_staged(sc, scenes.params(transfer=sc.transfer, tol=1e-6), steps, threads=min(16, os.cpu_count() or 1),
               counts_only=True)
