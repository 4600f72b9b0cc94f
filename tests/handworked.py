"""Hand-worked cases of the reference's rules that no closed form of the whole
stage covers: extrapolation's averaging, the top-row gravity wrap, the
particle-density splat and the density term of the right-hand side, the
sphere/box collision response of advection, the obstacle face speeds and
their serial order, the wall faces and the order of the wall passes, the
box push-out bounds and axis choice, and the order kept by sink removal.

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
DT_SMALL = 1e-4      # a stage time step for particles at rest


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


def _labels_faces(b, n, water, bg=(0.5, -0.25, 0.125)):
    """All AIR except the WATER cells listed; every face at the background."""
    lab = np.full((n, n, n), capi.AIR, np.uint8)
    for c in water:
        lab[c] = capi.WATER
    b.set_grid(capi.FIELD_LABEL, lab.reshape(-1))
    for f, v in zip(F1, bg):
        b.set_grid(f, np.full(n ** 3, v))
    return lab


def _face_changes(b, n, bg):
    """{(axis, cell): value} of every face that no longer holds the background."""
    out = {}
    for a, f in enumerate(F1):
        v = b.get_grid(f).reshape(n, n, n)
        for c in map(tuple, np.argwhere(v != np.float32(bg[a])).tolist()):
            out[(a, c)] = float(v[c])
    return out


def _face_triples(what, got: dict, expect: dict):
    """The set of changed faces must match exactly; their values to the bar."""
    keys = sorted(expect)
    return [(f"faces changed ({what})", sorted(got), keys),
            (f"values of the changed faces ({what})", np.array([got.get(k, np.nan) for k in keys]),
             np.array([expect[k] for k in keys], np.float64))]


def obstacle_face_speeds(make, sphere: bool):
    """addObstacle, macGrid.cpp:231-292, on an 8^3 grid (dx = 1/8).

    The obstacle covers the cells {3,4}^3: a box at the centre with edge 0.25
    (round(4 -+ 1) = 3 .. 5, upper bound exclusive, :268-271), or a sphere of
    radius 0.9 cells (0.1125) at the centre, whose cell centres (3.5 or 4.5)
    lie sqrt(0.75) = 0.87 cells away (:237-245; the next cells out are 1.66
    away).  Speed (0.3, -0.2, 0.1) for the box, (-0.3, 0.2, -0.1) for the
    sphere.  Labels are AIR except the WATER cells
      (2,3,3) -x of (3,3,3)   (5,4,4) +x of (4,4,4)   (3,2,4) -y of (3,3,4)
      (4,5,3) +y of (4,4,3)   (4,3,2) -z of (4,3,3)   (3,4,5) +z of (3,4,4)
      (4,4,3) inside the obstacle, (6,6,6) away from it;
    every face starts at (0.5, -0.25, 0.125).

    The cells are visited in x, y, z order (:241-243, :271-273).  Each
    becomes SOLID, writes its own face a with speed_a when its +a neighbour
    is WATER at that moment, and the -a neighbour's face a when that
    neighbour is WATER (:247-259, :275-287).  Worked through:
      (3,3,3): -x (2,3,3) WATER               -> u(2,3,3) = s_x
      (3,3,4): -y (3,2,4) WATER               -> v(3,2,4) = s_y
      (3,4,3): +x (4,4,3) still WATER         -> u(3,4,3) = s_x
      (3,4,4): +z (3,4,5) WATER               -> w(3,4,4) = s_z
      (4,3,3): +y (4,4,3) still WATER         -> v(4,3,3) = s_y
               -z (4,3,2) WATER               -> w(4,3,2) = s_z
      (4,3,4): nothing
      (4,4,3): now SOLID itself; +y (4,5,3)   -> v(4,4,3) = s_y
               (its -x and -y neighbours are already SOLID: no write)
      (4,4,4): +x (5,4,4) WATER               -> u(4,4,4) = s_x
    u(3,4,3) and v(4,3,3) are written only because (4,4,3) was still WATER
    when its lower neighbours were visited; a pass that first made all 8
    cells SOLID would leave them at the background."""
    n = 8
    grid = scenes.grid(n)
    b = make(grid)
    bg = (0.5, -0.25, 0.125)
    water = [(2, 3, 3), (5, 4, 4), (3, 2, 4), (4, 5, 3), (4, 3, 2), (3, 4, 5), (4, 4, 3), (6, 6, 6)]
    lab0 = _labels_faces(b, n, water, bg)
    if sphere:
        s = (-0.3, 0.2, -0.1)
        b.set_obstacles([capi.make_obstacle(capi.OBS_SPHERE, (0.5, 0.5, 0.5), speed=s, r=0.9 / n)])
    else:
        s = (0.3, -0.2, 0.1)
        b.set_obstacles([capi.make_obstacle(capi.OBS_BOX, (0.5, 0.5, 0.5), speed=s, size=(0.25, 0.25, 0.25))])
    b.run_stage(capi.STAGE_OBSTACLES)
    sx, sy, sz = s
    expect = {(0, (2, 3, 3)): sx, (1, (3, 2, 4)): sy, (0, (3, 4, 3)): sx, (2, (3, 4, 4)): sz, (1, (4, 3, 3)): sy,
              (2, (4, 3, 2)): sz, (1, (4, 4, 3)): sy, (0, (4, 4, 4)): sx}
    lab = b.get_grid(capi.FIELD_LABEL).reshape(n, n, n)
    lab_expect = lab0.copy()
    lab_expect[3:5, 3:5, 3:5] = capi.SOLID
    return _face_triples("obstacle speed", _face_changes(b, n, bg), expect) + [
        ("labels after the obstacle", int(np.count_nonzero(lab != lab_expect)), 0)]


def border_wall_faces(make, top_solid: int):
    """restoreBorderingSolidCellsAndSpeeds, macGrid.cpp:75-140, on an 8^3
    grid whose faces all start at (0.5, -0.25, 0.125).

    The release build runs the parallel branch (paralellDefine.h:5-7,
    macGrid.cpp:76-107): first the x-walls (x = 0 and 7 become SOLID; the face
    u(0,y,z) is zeroed when (1,y,z) is WATER and u(6,y,z) when (6,y,z) is
    WATER), then for each x the floor (y = 0 SOLID; v(x,0,z) zeroed when
    (x,1,z) is WATER), the top when isTopOfContainerSolid (y = 7 SOLID;
    v(x,6,z) zeroed when (x,6,z) is WATER), and after those the z-walls
    (z = 0 and 7 SOLID; w(x,y,0) zeroed when (x,y,1) is WATER, w(x,y,6) when
    (x,y,6) is WATER).  WATER cells and what they zero:
      (1,3,3) -> u(0,3,3)     (6,4,4) -> u(6,4,4)     (3,1,3) -> v(3,0,3)
      (4,6,4) -> v(4,6,4) only with a solid top      (3,4,1) -> w(3,4,0)
      (4,3,6) -> w(4,3,6)     (3,3,3) -> nothing (no wall next to it)
    and three WATER cells inside walls, which show the order of the passes:
      (0,1,5): an x-wall cell, already SOLID when the floor pass tests it:
               v(0,0,5) keeps its value;
      (5,0,1): a floor cell, already SOLID when the z-wall pass of x = 5
               tests it: w(5,0,0) keeps its value;
      (2,1,0): a z-wall cell the floor pass tests before the z-wall pass of
               x = 2 makes it SOLID: v(2,0,0) is zeroed."""
    n = 8
    grid = scenes.grid(n)
    b = make(grid, top_solid=top_solid)
    bg = (0.5, -0.25, 0.125)
    water = [(1, 3, 3), (6, 4, 4), (3, 1, 3), (4, 6, 4), (3, 4, 1), (4, 3, 6), (3, 3, 3), (0, 1, 5), (5, 0, 1),
             (2, 1, 0)]
    lab0 = _labels_faces(b, n, water, bg)
    b.run_stage(capi.STAGE_BORDERS)
    expect = {(0, (0, 3, 3)): 0.0, (0, (6, 4, 4)): 0.0, (1, (3, 0, 3)): 0.0, (2, (3, 4, 0)): 0.0,
              (2, (4, 3, 6)): 0.0, (1, (2, 0, 0)): 0.0}
    if top_solid:
        expect[(1, (4, 6, 4))] = 0.0
    lab = b.get_grid(capi.FIELD_LABEL).reshape(n, n, n)
    lab_expect = lab0.copy()
    lab_expect[0], lab_expect[-1], lab_expect[:, 0], lab_expect[:, :, 0], lab_expect[:, :, -1] = (capi.SOLID,) * 5
    if top_solid:
        lab_expect[:, -1] = capi.SOLID
    return _face_triples("wall", _face_changes(b, n, bg), expect) + [
        ("labels after the borders", int(np.count_nonzero(lab != lab_expect)), 0)]


def box_push_out(make):
    """pushParticlesOutOfObstacles for a box, simulator.cpp:279-310, on a
    16^3 grid (dx = 0.0625, particle radius r = 0.015625), box at the centre
    with edge 0.2.

    The bounds are not the box's own (0.4 .. 0.6) but the grid-aligned solid
    region of getMinMaxRect (macGrid.h:139-145): round(8 -+ 1.6) = 6 .. 10
    cells, 0.375 .. 0.625, widened by r to 0.359375 .. 0.640625 (:281-283).
    A particle inside those on every axis moves along the axis of least
    travel to the bound on its side of the box centre (the side test is
    pos > centre, :292); ties keep the first axis (strict <, :294,301):
      (0.62, 0.55, 0.5)   x: 0.020625 up, y: 0.090625, z: 0.140625 down -> x = 0.640625
      (0.45, 0.37, 0.52)  y: 0.010625 down is least                     -> y = 0.359375
      (0.5, 0.5, 0.6)     x, y: 0.140625 down (tie), z: 0.040625 up     -> z = 0.640625
      (0.6, 0.6, 0.45)    x, y: 0.040625 up (tie, x first), z: 0.090625 -> x = 0.640625
      (0.365, 0.5, 0.5)   outside the box + r (0.384375) but inside the
                          grid-aligned bounds: x 0.005625 down          -> x = 0.359375
      (0.3, 0.5, 0.5)     outside                                       -> unchanged
    None of them reaches the clamp box [dx + 1.01 r, 1 - dx - 1.01 r]."""
    grid = scenes.grid(16)
    b = make(grid)
    p0 = np.array([[0.62, 0.55, 0.5], [0.45, 0.37, 0.52], [0.5, 0.5, 0.6], [0.6, 0.6, 0.45], [0.365, 0.5, 0.5],
                   [0.3, 0.5, 0.5]])
    p0 = p0.astype(np.float32).astype(np.float64)
    b.set_particles(p0, np.zeros_like(p0))
    b.set_obstacles([capi.make_obstacle(capi.OBS_BOX, (0.5, 0.5, 0.5), size=(0.2, 0.2, 0.2))])
    b.run_stage(capi.STAGE_PUSH_OUT, DT_SMALL)
    p, _, _ = b.get_particles()
    expect = p0.copy()
    expect[0, 0] = 0.640625
    expect[1, 1] = 0.359375
    expect[2, 2] = 0.640625
    expect[3, 0] = 0.640625
    expect[4, 0] = 0.359375
    return [(f"particle {i} after the box push-out", p[i], expect[i]) for i in range(len(p0))]



def sink_removal_order(make):
    """Sinks with despawning on: advectParticles queues each particle whose
    centre ends within r_sink + r of the sink (simulator.cpp:182-193), and
    removeParticles drops the queued indices and shifts the survivors down
    in their old order (hashedParticles.cpp:158-172).  16^3 grid, sink of
    radius 0.1 at (0.5, 0.5, 0.5), so r_sink + r = 0.115625; 12 particles at
    rest at x = 0.3, 0.32, 0.34, 0.36, 0.38, 0.4, 0.42, 0.45, 0.48, 0.52,
    0.55, 0.7, y = 0.5, z = 0.5 + 0.001 i (particle i), so each is told apart
    by its z.  Particles 5 .. 10 (x = 0.4 .. 0.55) lie at most 0.1 from the
    centre and go; 0 .. 4 (x <= 0.38, 0.12 or more away) and 11 (0.2 away)
    stay, in that order."""
    grid = scenes.grid(16)
    b = make(grid, despawning_enabled=1)
    xs = [0.3, 0.32, 0.34, 0.36, 0.38, 0.4, 0.42, 0.45, 0.48, 0.52, 0.55, 0.7]
    p0 = np.array([[x, 0.5, 0.5 + 0.001 * i] for i, x in enumerate(xs)]).astype(np.float32).astype(np.float64)
    b.set_particles(p0, np.zeros_like(p0))
    b.set_obstacles([capi.make_obstacle(capi.OBS_SINK, (0.5, 0.5, 0.5), r=0.1)])
    b.run_stage(capi.STAGE_ADVECT, DT_SMALL)
    p, _, _ = b.get_particles()
    keep = [0, 1, 2, 3, 4, 11]
    return [("particle count after the sink", len(p), len(keep)), ("survivors in their old order", p, p0[keep])]


def solve_nan_break(make, dt: float):
    """solveIncompressibility's NaN exit, bridsonSolverGrid.cpp:262-292.

    8^3 grid, SOLID walls, WATER block 2..5 on every axis, all faces 0 except
    u2 = +inf on the x-face of cell (3,3,3).  The right-hand side then holds
    -inf at (3,3,3) and +inf at (4,3,3) (:79-89); its sum of squares is not
    below 1e-7 (:254-258), so the solve starts.  The MIC(0) sweep meets
    +inf - inf at (4,3,3) and z carries a NaN, so sigma = z.r and
    alpha = sigma / s.(A s) are NaN at the first iteration.  The loop breaks
    before multAdd (:270-272): 0 iterations, the pressure stays 0, and
    applying it (:291) leaves every face as it was."""
    n = 8
    grid = scenes.grid(n)
    b = make(grid)
    lab = np.full((n, n, n), capi.AIR, np.uint8)
    lab[0], lab[-1], lab[:, 0], lab[:, -1], lab[:, :, 0], lab[:, :, -1] = (capi.SOLID,) * 6
    lab[2:6, 2:6, 2:6] = capi.WATER
    b.set_grid(capi.FIELD_LABEL, lab.reshape(-1))
    u2 = np.zeros((n, n, n))
    u2[3, 3, 3] = np.inf
    b.set_grid(capi.FIELD_U2, u2.reshape(-1))
    b.set_grid(capi.FIELD_V2, np.zeros(grid.nc))
    b.set_grid(capi.FIELD_W2, np.zeros(grid.nc))
    b.run_stage(capi.STAGE_SOLVE, dt)
    st = b.stats()
    rhs = b.get_grid(capi.FIELD_RHS).reshape(n, n, n)
    faces = [b.get_grid(f).reshape(n, n, n) for f in F2]
    return [("iterations", st.pcg_iterations, 0), ("NaN exit reported", st.solve_nan, 1),
            ("early exit", st.early_exit, 0),
            ("rhs at (3,3,3) and (4,3,3)", [float(rhs[3, 3, 3]), float(rhs[4, 3, 3])], [-np.inf, np.inf]),
            ("pressure", float(np.abs(b.get_grid(capi.FIELD_PRESSURE)).max()), 0.0),
            ("u2 after the solve", [bool(np.array_equal(faces[0], u2))], [True]),
            ("v2, w2 after the solve", float(max(np.abs(faces[1]).max(), np.abs(faces[2]).max())), 0.0)]


def gfx_snapshot_linear_density(make):
    """SimulationManager's gfx snapshot, simulationManager.cpp:223-231, on a
    16^3 grid whose density field is linear in the cell centres,
    avgPNum(x, y, z) = 2 + 0.5 x - 0.25 y + 0.125 z (cell units).

    The density of a particle sums, over the 8 cells of getCellsAround
    (macGrid.cpp:188-202), the tent weight of the cell centre times its
    avgPNum.  The 8 weights are the trilinear ones, so a linear field is
    reproduced exactly: the density is 2 + 0.5 (px/dx - 0.5) - 0.25
    (py/dx - 0.5) + 0.125 (pz/dx - 0.5).  |v| is glm::length(v): particle
    velocities (3, 4, 12) give 13, (0, -5, 0) give 5.  Records come out in
    the particle order."""
    n = 16
    grid = scenes.grid(n)
    dx = grid.cell_d
    b = make(grid)
    i = np.arange(n, dtype=np.float64)
    X, Y, Z = np.meshgrid(i, i, i, indexing="ij")
    b.set_grid(capi.FIELD_DENSITY, (2.0 + 0.5 * X - 0.25 * Y + 0.125 * Z).reshape(-1))
    pos = np.array([[3.3, 5.7, 8.1], [10.5, 2.25, 4.75], [7.0, 7.0, 7.0], [1.6, 13.9, 12.2]]) * dx
    pos = pos.astype(np.float32).astype(np.float64)
    vel = np.array([[3.0, 4.0, 12.0], [0.0, -5.0, 0.0], [0.0, 0.0, 0.0], [3.0, 4.0, 12.0]])
    b.set_particles(pos, vel)
    snap = b.gfx_snapshot()
    if snap.dtype.names:   # the HIP library's records -> (n, 5)
        snap = np.column_stack([snap["pos"], snap["v"], snap["density"]]).astype(np.float64)
    g = pos / dx - 0.5
    expect_d = 2.0 + 0.5 * g[:, 0] - 0.25 * g[:, 1] + 0.125 * g[:, 2]
    return [("positions in particle order", snap[:, :3], pos), ("speeds", snap[:, 3], np.array([13.0, 5.0, 0.0, 13.0])),
            ("densities", snap[:, 4], expect_d)]


OMEGA_SOR = 1.98   # basicMacGrid.cpp:13


def basic_sor(make, iterations: int, pressure_enabled: int = 0):
    """BasicMacGrid::solveIncompressibility, basicMacGrid.cpp:5-102, release
    (parallel) branch: each iteration a pass over the WATER cells with
    x + y + z odd, then one over the even ones; a cell with s non-SOLID
    neighbours computes d = (-u - v - w + u(-x) + v(-y) + w(-z)
    [+ (avgPNum - averagePressure) pressureK]) * 1.98 / s and adds d to its own
    three faces and subtracts it from the three faces it shares with its -x,
    -y, -z neighbours, each only where the neighbour across it is not SOLID.
    8^3 grid, SOLID walls, faces 0 except as listed.

    Cell A (3,3,3), isolated WATER (all 6 neighbours AIR), odd.  Its faces:
    u(3,3,3) = 0.3, v(3,3,3) = -0.2, w(3,3,3) = 0.1, u(2,3,3) = -0.5,
    v(3,2,3) = 0.25, w(3,3,2) = 0.05: the bracket D0 = -0.4 (+ the density
    term T).  Each update changes the bracket by -6d = -1.98 D, so after k
    iterations the own faces have gained 0.33 D0 (1 - (-0.98)^k) / 1.98 and
    the -side faces lost as much.

    Cells B (2,5,5), even, and C (3,5,5), odd, WATER side by side; their
    shared face u(2,5,5) = 0.6 and u(3,5,5) = 0.2.  The odd pass goes first:
    C sees -0.2 + 0.6 = 0.4, d_C = 0.132 (u(3,5,5) 0.332, v, w of C 0.132,
    shared face 0.468, v(3,4,5) and w(3,5,4) -0.132); then B sees -0.468,
    d_B = -0.15444 (shared face 0.31356, v, w of B -0.15444, u(1,5,5),
    v(2,4,5), w(2,5,4) 0.15444).  (A lexicographic sweep would update B
    first.)  One iteration for this pair.

    Cell W (1,2,6), odd, WATER next to the x = 0 and z = 7 walls: s = 4,
    u(1,2,6) = 0.5 gives d = -0.5 * 1.98 / 4 = -0.2475; its faces on the
    walls, u(0,2,6) and w(1,2,6), stay 0."""
    n = 8
    grid = scenes.grid(n)
    kw = dict(precond=capi.SOLVER_BASIC_SOR, max_iterations=iterations, pressure_enabled=pressure_enabled,
              pressure_k=2.0, average_pressure=2.0)
    b = make(grid, **kw)
    lab = np.full((n, n, n), capi.AIR, np.uint8)
    lab[0], lab[-1], lab[:, 0], lab[:, -1], lab[:, :, 0], lab[:, :, -1] = (capi.SOLID,) * 6
    for c in ((3, 3, 3), (2, 5, 5), (3, 5, 5), (1, 2, 6)):
        lab[c] = capi.WATER
    b.set_grid(capi.FIELD_LABEL, lab.reshape(-1))
    f = [np.zeros((n, n, n)) for _ in range(3)]
    f[0][3, 3, 3], f[1][3, 3, 3], f[2][3, 3, 3] = 0.3, -0.2, 0.1
    f[0][2, 3, 3], f[1][3, 2, 3], f[2][3, 3, 2] = -0.5, 0.25, 0.05
    f[0][2, 5, 5], f[0][3, 5, 5] = 0.6, 0.2
    f[0][1, 2, 6] = 0.5
    for fld, v in zip(F2, f):
        b.set_grid(fld, v.reshape(-1))
    dens = np.zeros((n, n, n))
    dens[3, 3, 3] = 2.5   # T = (2.5 - 2) * 2 = 1 at A when the density term is on; B, C, W: (0 - 2) * 2 = -4
    b.set_grid(capi.FIELD_DENSITY, dens.reshape(-1))
    b.run_stage(capi.STAGE_SOLVE, scenes.DT)
    out = [b.get_grid(fld).reshape(n, n, n) for fld in F2]
    res = [("iterations reported", b.stats().pcg_iterations, iterations)]
    # A: closed form
    d0 = -0.4 + (1.0 if pressure_enabled else 0.0)
    gain = 0.33 * d0 * (1.0 - (-0.98) ** iterations) / 1.98
    res += [("A own faces", [out[0][3, 3, 3], out[1][3, 3, 3], out[2][3, 3, 3]], [0.3 + gain, -0.2 + gain, 0.1 + gain]),
            ("A -side faces", [out[0][2, 3, 3], out[1][3, 2, 3], out[2][3, 3, 2]],
             [-0.5 - gain, 0.25 - gain, 0.05 - gain])]
    if iterations == 1 and not pressure_enabled:
        res += [("C own faces", [out[0][3, 5, 5], out[1][3, 5, 5], out[2][3, 5, 5]], [0.332, 0.132, 0.132]),
                ("C -y, -z faces", [out[1][3, 4, 5], out[2][3, 5, 4]], [-0.132, -0.132]),
                ("shared face", out[0][2, 5, 5], 0.468 - 0.15444),
                ("B v, w", [out[1][2, 5, 5], out[2][2, 5, 5]], [-0.15444, -0.15444]),
                ("B -side faces", [out[0][1, 5, 5], out[1][2, 4, 5], out[2][2, 5, 4]], [0.15444, 0.15444, 0.15444]),
                ("W own faces", [out[0][1, 2, 6], out[1][1, 2, 6], out[2][1, 2, 6]], [0.2525, -0.2475, 0.0]),
                ("W -side faces", [out[0][0, 2, 6], out[1][1, 1, 6], out[2][1, 2, 5]], [0.0, 0.2475, 0.2475])]
    return [(what, np.asarray(got, np.float64), np.asarray(exp, np.float64)) for what, got, exp in res]


def push_apart_pair(make):
    """pushParticlesApart, hashedParticles.cpp:64-107, on a 16^3 grid
    (dx = 0.0625, r = 0.015625, diameter D = 0.03125):
      P0 (0.5, 0.5, 0.5) and P1 (0.52, 0.5, 0.5), 0.02 apart: each moves half
        the overlap, 0.005625, away from the other (:91-95), so they end
        D apart with the midpoint kept: x = 0.494375 and 0.525625;
      P2 (0.3, 0.3, 0.3) and P3 (0.3, 0.34, 0.3), 0.04 > D apart: unchanged
        (:88);
      P4 and P5 at (0.7, 0.7, 0.7), 0.00005 apart (closer than 1e-4): left as
        they are (:88)."""
    grid = scenes.grid(16)
    b = make(grid, push_apart_enabled=1)
    p0 = np.array([[0.5, 0.5, 0.5], [0.52, 0.5, 0.5], [0.3, 0.3, 0.3], [0.3, 0.34, 0.3], [0.7, 0.7, 0.7],
                   [0.70005, 0.7, 0.7]]).astype(np.float32).astype(np.float64)
    b.set_particles(p0, np.zeros_like(p0))
    b.run_stage(capi.STAGE_PUSH_APART)
    p, _, _ = b.get_particles()
    half = (2 * grid.particle_r - (p0[1, 0] - p0[0, 0])) / 2
    expect = p0.copy()
    expect[0, 0] -= half
    expect[1, 0] += half
    return [("pair pushed to one diameter", p[:2], expect[:2]), ("pair further than D", p[2:4], p0[2:4]),
            ("coincident pair", p[4:], p0[4:]),
            ("distance after", float(np.linalg.norm(p[1] - p[0])), 2 * grid.particle_r)]
