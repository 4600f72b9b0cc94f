"""Hand-worked cases of the reference's rules that no closed form of the whole
stage covers: extrapolation's averaging, the top-row gravity wrap, the
particle-density splat and the density term of the right-hand side, and the
sphere/box collision response of advection.

The expected numbers are worked out by hand from the reference's text (cited
per case) on inputs small enough to follow cell by cell; the collision cases
use geometry instead (the exit time of a ray from a sphere or a slab, and
specular reflection).  None of them re-runs the oracle's code.  Each case
takes `make(grid, **params)`, which returns a backend with the library's
stage interface (set_grid / get_grid / run_stage / set_particles /
get_particles / set_obstacles): the FP64 oracle, or the HIP library through
fluid_simulator_amd.capi.  It returns (what, got, expected) triples.
"""
from __future__ import annotations

import math

import numpy as np

from fluid_simulator_amd import capi, scenes

F1 = (capi.FIELD_U, capi.FIELD_V, capi.FIELD_W)
F2 = (capi.FIELD_U2, capi.FIELD_V2, capi.FIELD_W2)
BACK_STEP = 0.0002   # simulator.cpp:197,219 (seconds per back-step)


def extrapolation_two_layers(make):
    """extrapolateVelocities, macGrid.cpp:294-336, on a 7^3 grid with three
    WATER cells A (1,3,3), B (3,3,3), C (5,3,3) on one line.

    Rules read from the reference: a cell that is not yet valid takes, for
    each component, the mean over its valid 6-neighbours of their faces of
    that component (:307-324); it writes face a only when its +a neighbour is
    inside the grid and not WATER (:326-331); cells made valid in pass `it`
    count from pass it+1 on (valid <= it, :297-299); two passes (:295).
    Background faces (0.5, -0.25, 0.125); A = (1, 2, 3), B = (5, 7, 11),
    C = (13, 17, 19)."""
    n = 7
    grid = scenes.grid(n)
    b = make(grid)
    bg = (0.5, -0.25, 0.125)
    lab = np.full((n, n, n), capi.AIR, np.uint8)
    fields = [np.full((n, n, n), c) for c in bg]
    for cell, vals in {(1, 3, 3): (1.0, 2.0, 3.0), (3, 3, 3): (5.0, 7.0, 11.0), (5, 3, 3): (13.0, 17.0, 19.0)}.items():
        lab[cell] = capi.WATER
        for a in range(3):
            fields[a][cell] = vals[a]
    b.set_grid(capi.FIELD_LABEL, lab.reshape(-1))
    for f, v in zip(F2, fields):
        b.set_grid(f, v.reshape(-1))
    b.run_stage(capi.STAGE_EXTRAPOLATE)
    out = [b.get_grid(f).reshape(n, n, n) for f in F2]
    expect = {
        (1, 3, 3): (1.0, 2.0, 3.0),        # WATER: a source, never written
        # pass 0: the mean of A and B; the +x neighbour is B (WATER), so u keeps the background
        (2, 3, 3): (0.5, 4.5, 7.0),
        (4, 3, 3): (0.5, 12.0, 15.0),      # mean of B and C, +x neighbour C
        (6, 3, 3): (0.5, 17.0, 19.0),      # C alone; the +x neighbour is outside the grid
        (0, 3, 3): (0.5, 2.0, 3.0),        # A alone; +x neighbour A
        (1, 4, 3): (1.0, 2.0, 3.0),        # A alone, every face written
        (3, 4, 3): (5.0, 7.0, 11.0),
        # pass 1: valid neighbours (1,4,3), (3,4,3) and (2,3,3), whose u kept 0.5
        (2, 4, 3): ((1.0 + 5.0 + 0.5) / 3.0, (2.0 + 7.0 + 4.5) / 3.0, (3.0 + 11.0 + 7.0) / 3.0),
        # pass 1 only: (0,3,3) and (1,4,3) became valid in pass 0, which pass 0 itself does not see
        (0, 4, 3): ((0.5 + 1.0) / 2.0, 2.0, 3.0),
        (3, 6, 3): bg,                     # three cells from the water: untouched
        (6, 6, 6): bg,
    }
    return [(f"extrapolated faces of cell {c}", tuple(float(o[c]) for o in out), e) for c, e in expect.items()]


def post_p2g_gravity_wrap(make, dt: float):
    """postP2GUpdate, macGrid.cpp:204-219: v2 = v, and v2_y += g dt iff the
    cell and cell<1,1>(pos) are not SOLID (:209-210).  cell<1,1> is the cell
    one flat index row further, so for the top row y = N-1 it is (x+1, 0, z),
    a SOLID floor cell: the top row gets no gravity.  8^3 grid, SOLID walls
    x = 0, x = 7, y = 0, z = 0, z = 7 (the top is open), one SOLID cell
    (3,4,3); v = 0.25, u = 0.5, w = -0.125 everywhere."""
    n = 8
    grid = scenes.grid(n)
    prm = scenes.params(gravity_enabled=1)
    b = make(grid, gravity_enabled=1)
    lab = np.full((n, n, n), capi.AIR, np.uint8)
    lab[0], lab[-1], lab[:, 0], lab[:, :, 0], lab[:, :, -1] = (capi.SOLID,) * 5
    lab[3, 4, 3] = capi.SOLID
    b.set_grid(capi.FIELD_LABEL, lab.reshape(-1))
    for f, c in zip(F1, (0.5, 0.25, -0.125)):
        b.set_grid(f, np.full(grid.nc, c))
    b.run_stage(capi.STAGE_POST_P2G, dt)
    u2, v2, w2 = (b.get_grid(f).reshape(n, n, n) for f in F2)
    g = float(np.float32(prm.gravity)) * dt
    expect = {
        (3, 5, 3): 0.25 + g,   # +y neighbour AIR
        (3, 6, 3): 0.25 + g,   # +y neighbour is the open top row
        (3, 7, 3): 0.25,       # top row: the wrapped neighbour (4, 0, 3) is the SOLID floor
        (6, 7, 5): 0.25,       # top row next to the x wall: neighbour (7, 0, 5)
        (3, 3, 3): 0.25,       # below the SOLID cell
        (3, 4, 3): 0.25,       # SOLID itself
        (3, 0, 3): 0.25,       # floor
        (2, 1, 2): 0.25 + g,
    }
    res = [(f"v2 of cell {c}", float(v2[c]), e) for c, e in expect.items()]
    res.append(("u2 copied from u", float(np.abs(u2 - 0.5).max()), 0.0))
    res.append(("w2 copied from w", float(np.abs(w2 + 0.125).max()), 0.0))
    return res


def mark_density_splat(make):
    """markFluidCellsAndCalculateParticleDensities, simulator.cpp:354-368:
    the particle's cell (truncated pos/dx) becomes WATER, and each of the 8
    cells of getCellsAround (macGrid.cpp:188-202: base = trunc(pos/dx - 0.5))
    gains the tent weight of its centre (x + 0.5) dx (macGrid.cpp:20-33).
    8^3 grid, dx = 1/8, two particles (cell units):
    P1 (2.75, 2.5, 2.5): weights 0.75 to (2,2,2) and 0.25 to (3,2,2);
    P2 (4.5, 5.0, 5.0): 0.25 to each of (4,4,4), (4,5,4), (4,4,5), (4,5,5)."""
    n = 8
    grid = scenes.grid(n)
    dx = grid.cell_d
    b = make(grid, pressure_enabled=1)
    b.set_particles(np.array([[2.75 * dx, 2.5 * dx, 2.5 * dx], [4.5 * dx, 5.0 * dx, 5.0 * dx]]))
    b.run_stage(capi.STAGE_P2G)
    b.run_stage(capi.STAGE_MARK)
    dens = b.get_grid(capi.FIELD_DENSITY).reshape(n, n, n)
    lab = b.get_grid(capi.FIELD_LABEL).reshape(n, n, n)
    expect = {(2, 2, 2): 0.75, (3, 2, 2): 0.25, (4, 4, 4): 0.25, (4, 5, 4): 0.25, (4, 4, 5): 0.25, (4, 5, 5): 0.25}
    res = [(f"density of cell {c}", float(dens[c]), e) for c, e in expect.items()]
    rest = dens.copy()
    for c in expect:
        rest[c] = 0.0
    res.append(("density elsewhere", float(np.abs(rest).max()), 0.0))
    water = sorted(map(tuple, np.argwhere(lab == capi.WATER).tolist()))
    res.append(("WATER cells", water, [(2, 2, 2), (4, 5, 5)]))
    return res


def rhs_density_term(make, dt: float, pressure_enabled: int):
    """calculateRHS, bridsonSolverGrid.cpp:79-89:
    b = -(1/dx)(u2 + v2 + w2 - u2(-x) - v2(-y) - w2(-z)) + [pressureEnabled] (avgPNum - averagePressure) pressureK.
    8^3 grid (dx = 1/8), SOLID walls, WATER block 2..5 on every axis,
    u2 = 0.125 x, v2 = w2 = 0 (divergence 0.125 per cell, so the first term is
    -1), density 3.0 at (3,3,3) and 1.5 at (4,4,4), K = 2, averagePressure = 2."""
    n = 8
    grid = scenes.grid(n)
    b = make(grid, pressure_enabled=pressure_enabled, pressure_k=2.0, average_pressure=2.0)
    lab = np.full((n, n, n), capi.AIR, np.uint8)
    lab[0], lab[-1], lab[:, 0], lab[:, -1], lab[:, :, 0], lab[:, :, -1] = (capi.SOLID,) * 6
    lab[2:6, 2:6, 2:6] = capi.WATER
    b.set_grid(capi.FIELD_LABEL, lab.reshape(-1))
    u2 = np.broadcast_to(0.125 * np.arange(n)[:, None, None], (n, n, n))
    b.set_grid(capi.FIELD_U2, u2.reshape(-1))
    b.set_grid(capi.FIELD_V2, np.zeros(grid.nc))
    b.set_grid(capi.FIELD_W2, np.zeros(grid.nc))
    dens = np.zeros((n, n, n))
    dens[3, 3, 3], dens[4, 4, 4] = 3.0, 1.5
    b.set_grid(capi.FIELD_DENSITY, dens.reshape(-1))
    b.run_stage(capi.STAGE_SOLVE, dt)
    rhs = b.get_grid(capi.FIELD_RHS).reshape(n, n, n)
    if pressure_enabled:
        expect = {(3, 3, 3): -1.0 + (3.0 - 2.0) * 2.0, (4, 4, 4): -1.0 + (1.5 - 2.0) * 2.0,
                  (2, 2, 2): -1.0 + (0.0 - 2.0) * 2.0, (5, 2, 4): -5.0, (6, 6, 6): 0.0, (1, 3, 3): 0.0}
    else:
        expect = {(3, 3, 3): -1.0, (4, 4, 4): -1.0, (2, 2, 2): -1.0, (5, 2, 4): -1.0, (6, 6, 6): 0.0}
    return [(f"rhs of cell {c}", float(rhs[c]), e) for c, e in expect.items()]


def _sphere_exit_time(d, v, R):
    """Smallest tau >= 0 with |d - tau v| = R for |d| < R (the point stepped back along -v leaves the sphere)."""
    a, dv, c = float(v @ v), float(d @ v), float(d @ d - R * R)
    return (dv + math.sqrt(dv * dv - a * c)) / a


def sphere_bounce(make, oblique: bool, speed_x: float = 0.0):
    """advectParticles' sphere response, simulator.cpp:178-213, for a sphere
    (r 0.1, now at the centre of a 16^3 grid, moving at speed_x along x) and
    one particle at 10 m/s along +x for dt = 0.01 that ends inside it (no wall
    within reach).  The reference steps the particle back along -v and the
    sphere along -speed in 0.0002 s steps while they overlap (:196-197),
    takes one more step (:200), and reflects the velocity relative to the
    sphere about the normal from the stepped-back sphere centre, restitution
    1 (:201-208); the particle then moves on for the time it stepped back
    (:209).  Expected from geometry: k = ceil(tau / 0.0002) for the exit time
    tau of the relative motion, back-step (k+1) 0.0002 s, velocity
    w - 2 (w.n) n + speed with w = v - speed."""
    grid = scenes.grid(16)
    c = np.array([0.5, 0.5, 0.5])
    s = np.array([speed_x, 0.0, 0.0])
    R = 0.1 + grid.particle_r
    v = np.array([10.0, 0.0, 0.0])
    w = v - s
    dt = 0.01
    p_end = c + (np.array([-0.108, 0.03, 0.0]) if oblique else np.array([-(R - 2.5 * BACK_STEP * w[0]), 0.0, 0.0]))
    p0 = p_end - v * dt
    b = make(grid)
    b.set_particles(p0[None, :], v[None, :])
    b.set_obstacles([capi.make_obstacle(capi.OBS_SPHERE, tuple(c), speed=tuple(s), r=0.1)])
    b.run_stage(capi.STAGE_ADVECT, dt)
    p, vel, _ = b.get_particles()
    tau = _sphere_exit_time(p_end - c, w, R)
    steps = tau / BACK_STEP
    assert 0.1 < steps - math.floor(steps) < 0.9, "the case must not sit on a back-step boundary"
    assert steps < 4.5, "the case must stay below the back-step cap"
    back = (math.ceil(steps) + 1) * BACK_STEP
    q = p_end - back * v
    centre = c - s * back
    nrm = (q - centre) / np.linalg.norm(q - centre)
    v1 = w - 2.0 * float(w @ nrm) * nrm + s
    p1 = q + back * v1
    assert np.linalg.norm(p1 - c) > R, "the case must end with one bounce"
    return [("position after the bounce", p[0], p1), ("velocity after the bounce", vel[0], v1),
            ("particle count", len(p), 1)]


def sphere_back_step_cap(make):
    """The back-step loop of simulator.cpp:196 also stops once backTime is no
    longer below t + 0.001f (a float literal, 0.0010000000475).  Five steps of
    0.0002 s add up to 0.001 in double, which is still below it, so a first
    collision (t = 0) backs off at most six steps, plus the extra step (:200):
    7 x 0.0002 s.  Case: a static sphere (r 0.1) and a particle at 10 m/s
    along +x ending 0.02 deep, which would need 10 steps to get out; after
    backing off 0.014 it is still 0.006 inside, the normal is -x, its
    velocity reverses, and the remaining 0.0014 s carry it 0.014 further out
    (0.008 outside)."""
    grid = scenes.grid(16)
    c = np.array([0.5, 0.5, 0.5])
    R = 0.1 + grid.particle_r
    v = np.array([10.0, 0.0, 0.0])
    dt = 0.01
    p_end = c - np.array([R - 0.02, 0.0, 0.0])
    b = make(grid)
    b.set_particles((p_end - v * dt)[None, :], v[None, :])
    b.set_obstacles([capi.make_obstacle(capi.OBS_SPHERE, tuple(c), r=0.1)])
    b.run_stage(capi.STAGE_ADVECT, dt)
    p, vel, _ = b.get_particles()
    back = 7 * BACK_STEP
    v1 = -v
    p1 = p_end - back * v + back * v1
    return [("position after the capped bounce", p[0], p1), ("velocity after the capped bounce", vel[0], v1)]


def moving_box_bounce(make):
    """advectParticles' box response, simulator.cpp:214-240, for a box (size
    0.2, centre 0.5) moving at +1 m/s in x, and one particle at 10 m/s along
    +x for dt = 0.01 that ends inside its low-x face, on a 16^3 grid.
    Stepping back, the particle and the box close at 9 m/s, so it leaves the
    box after ceil(pen / (9 * 0.0002)) steps, pen being its depth; one more
    step follows (:223).  The low-x face branch (:231-233) sets
    v_x = -v_x * 0.2 - speed_x = -3 (the reference subtracts the box speed on
    that face), and the particle moves on for the time it stepped back."""
    grid = scenes.grid(16)
    rp = grid.particle_r
    c = np.array([0.5, 0.5, 0.5])
    speed = np.array([1.0, 0.0, 0.0])
    v = np.array([10.0, 0.0, 0.0])
    dt = 0.01
    lower = c[0] - 0.1 - rp
    pen = 2.5 * BACK_STEP * (v[0] - speed[0])
    p_end = np.array([lower + pen, 0.5, 0.5])
    b = make(grid)
    b.set_particles((p_end - v * dt)[None, :], v[None, :])
    b.set_obstacles([capi.make_obstacle(capi.OBS_BOX, tuple(c), speed=tuple(speed), size=(0.2, 0.2, 0.2))])
    b.run_stage(capi.STAGE_ADVECT, dt)
    p, vel, _ = b.get_particles()
    back = (math.ceil(pen / (BACK_STEP * (v[0] - speed[0]))) + 1) * BACK_STEP
    v1 = np.array([-v[0] * 0.2 - speed[0], 0.0, 0.0])
    p1 = p_end - back * v + back * v1
    return [("position after the bounce", p[0], p1), ("velocity after the bounce", vel[0], v1)]
