"""The hand-worked cases of tests/handworked.py on the HIP library: the same
inputs and the same hand-derived expectations as the oracle's pins, so the
GPU path is held to the reference's rules directly, not only through the
oracle.  Bars: FP32 fields within 2 ulps of the expected value's scale,
particle positions and velocities within 2e-6 + 1e-6 relative (FP32
storage of positions near 0.4 and velocities near 10)."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import handworked
from fluid_simulator_amd import capi, scenes

pytestmark = pytest.mark.gpu
DT = scenes.DT


def say(msg: str):
    print(msg, flush=True)
    sys.stdout.flush()


def make(grid, **kw):
    return capi.Simulation(grid, scenes.params(**kw))


def _check(triples, rtol, atol):
    for what, got, expect in triples:
        if isinstance(expect, list):
            assert got == expect, what
        else:
            np.testing.assert_allclose(np.asarray(got, float), np.asarray(expect, float), rtol=rtol, atol=atol,
                                       err_msg=what)


def test_extrapolation_hand_worked(gpu):
    say("case: extrapolation, 7^3, three WATER cells")
    _check(handworked.extrapolation_two_layers(make), 2.4e-7, 0.0)


def test_top_row_gravity_wrap_hand_worked(gpu):
    say("case: post-P2G gravity with the top-row wrap, 8^3")
    _check(handworked.post_p2g_gravity_wrap(make, DT), 2.4e-7, 0.0)


def test_density_splat_hand_worked(gpu):
    say("case: particle density splat, 8^3, two particles")
    _check(handworked.mark_density_splat(make), 0.0, 1e-7)


@pytest.mark.parametrize("pressure_enabled", [1, 0])
def test_rhs_density_term_hand_worked(gpu, pressure_enabled):
    say(f"case: RHS with density term {pressure_enabled}, 8^3")
    _check(handworked.rhs_density_term(make, DT, pressure_enabled), 2.4e-7, 1e-7)


@pytest.mark.parametrize("oblique", [False, True])
def test_sphere_bounce_geometry(gpu, oblique):
    say(f"case: sphere bounce, oblique {oblique}")
    _check(handworked.sphere_bounce(make, oblique), 1e-6, 2e-6)


def test_moving_box_bounce_hand_worked(gpu):
    say("case: moving box bounce")
    _check(handworked.moving_box_bounce(make), 1e-6, 2e-6)


@pytest.mark.parametrize("oblique", [False, True])
def test_moving_sphere_bounce_geometry(gpu, oblique):
    say(f"case: moving sphere bounce, oblique {oblique}")
    _check(handworked.sphere_bounce(make, oblique, speed_x=-2.0), 1e-6, 2e-6)


def test_sphere_back_step_cap_hand_worked(gpu):
    say("case: sphere bounce at the back-step cap")
    _check(handworked.sphere_back_step_cap(make), 1e-6, 2e-6)


@pytest.mark.parametrize("sphere", [False, True])
def test_obstacle_face_speeds_hand_worked(gpu, sphere):
    say(f"case: obstacle face speeds, 8^3, sphere {sphere}")
    _check(handworked.obstacle_face_speeds(make, sphere), 2.4e-7, 0.0)


@pytest.mark.parametrize("top_solid", [0, 1])
def test_border_wall_faces_hand_worked(gpu, top_solid):
    say(f"case: wall faces, 8^3, top solid {top_solid}")
    _check(handworked.border_wall_faces(make, top_solid), 0.0, 0.0)


def test_box_push_out_hand_worked(gpu):
    say("case: box push-out, 16^3")
    _check(handworked.box_push_out(make), 0.0, 0.0)


def test_sink_removal_order_hand_worked(gpu):
    say("case: sink removal keeps the survivors' order, 16^3")
    _check(handworked.sink_removal_order(make), 0.0, 0.0)


def test_solve_nan_break_hand_worked(gpu):
    say("case: pressure solve ends on a NaN alpha (an infinite face), 8^3")
    _check(handworked.solve_nan_break(make, DT), 0.0, 0.0)


def test_gfx_snapshot_linear_density_hand_worked(gpu):
    say("case: gfx snapshot of a linear density field, 16^3")
    _check(handworked.gfx_snapshot_linear_density(make), 4e-6, 4e-6)


@pytest.mark.parametrize("iterations,pressure_enabled", [(1, 0), (5, 0), (4, 1)])
def test_basic_sor_hand_worked(gpu, iterations, pressure_enabled):
    say(f"case: BasicMacGrid red-black SOR, {iterations} iterations, density term {pressure_enabled}, 8^3")
    _check(handworked.basic_sor(make, iterations, pressure_enabled), 2.4e-7, 1e-7)


def test_push_apart_pair_hand_worked(gpu):
    say("case: push-apart of a pair, 16^3")
    _check(handworked.push_apart_pair(make), 0.0, 1e-7)
