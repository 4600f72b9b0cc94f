"""Pins of the FP64 oracle's 2D mode (CPU only).

The 2D mode (macGrid.cpp:9-13) is an n x n x 3 grid whose cells are
cellD.z = depth / 3 deep, with particles held at z = depth / 2.  What
changes against 3D is per-axis arithmetic (the z cell size in positions,
weights, bin keys and rasterisation), the z clamp without the particle
radius (simulator.cpp:140), v.z = 0 in G2P (simulator.cpp:382-385) and the
fixed z of push-apart (hashedParticles.cpp:96-101).  Each is pinned by what
the geometry fixes: affine fields reproduced by trilinear weights, closed
forms, brute force, and the 5-point Laplacian the middle layer reduces to.
"""
from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle
from fluid_simulator_amd import capi, scenes
from test_oracle_pins import laplacian

DT = scenes.DT
DEPTH = scenes.DEPTH_2D


def mk(grid, **kw):
    return oracle.Oracle(grid, scenes.params(**kw))


def rand_2d(grid, n, seed, z_any=False):
    """Positions inside the clamp box; z = depth / 2 unless z_any (then
    anywhere in the middle layer)."""
    rng = np.random.default_rng(seed)
    lo = grid.cell_d + 1.01 * grid.particle_r
    pos = rng.uniform(lo, grid.nx * grid.cell_d - lo, (n, 3))
    dz = grid.cell_dz
    pos[:, 2] = rng.uniform(dz, 2 * dz, n) if z_any else DEPTH / 2
    return pos


def face_positions(grid):
    i, k = np.arange(grid.nx), np.arange(3)
    X, Y, Z = np.meshgrid(i, i, k, indexing="ij")
    d, dz = grid.cell_d, grid.cell_dz
    return [np.stack([(X + 1) * d, (Y + .5) * d, (Z + .5) * dz], -1).reshape(-1, 3),
            np.stack([(X + .5) * d, (Y + 1) * d, (Z + .5) * dz], -1).reshape(-1, 3),
            np.stack([(X + .5) * d, (Y + .5) * d, (Z + 1) * dz], -1).reshape(-1, 3)]


V0 = np.array([1.0, -0.5, 0.25])
G = np.array([[0.5, -1.0, 0.25], [2.0, 0.1, -0.3], [-0.7, 0.4, 1.5]])


def test_2d_grid_description_rules():
    """2D needs nz = 3 and cell_d_z > 0; 3D accepts cell_d_z = 0 or = cell_d only."""
    ok = scenes.grid_2d(8)
    mk(ok).close()
    for bad in (capi.Grid(8, 8, 4, 1 / 8, 1 / 32, two_d=1, cell_d_z=0.03),
                capi.Grid(8, 8, 3, 1 / 8, 1 / 32, two_d=1, cell_d_z=0.0),
                capi.Grid(8, 8, 8, 1 / 8, 1 / 32, two_d=0, cell_d_z=0.03)):
        with pytest.raises(RuntimeError):
            mk(bad)
    mk(capi.Grid(8, 8, 8, 1 / 8, 1 / 32, two_d=0, cell_d_z=1 / 8)).close()


@pytest.mark.parametrize("transfer", [capi.PIC, capi.FLIP, capi.APIC])
def test_2d_g2p_affine_field_anisotropic_cells(transfer):
    """Trilinear weights with cellD.z = depth / 3 reproduce an affine face field exactly (and its
    gradient for APIC) on x and y; v.z is set to 0 and the z row of C is left as it was."""
    grid = scenes.grid_2d(12)
    o = mk(grid, transfer=transfer, flip_ratio=0.0)
    pos = rand_2d(grid, 3000, 3, z_any=True)
    c_in = np.full((len(pos), 9), 7.0)
    o.set_particles(pos, np.full_like(pos, 3.0), c_in)
    fp = face_positions(grid)
    for a, (f, f1) in enumerate(((capi.FIELD_U2, capi.FIELD_U), (capi.FIELD_V2, capi.FIELD_V),
                                  (capi.FIELD_W2, capi.FIELD_W))):
        o.set_grid(f, V0[a] + fp[a] @ G[a])
        o.set_grid(f1, V0[a] + fp[a] @ G[a])
    o.run_stage(capi.STAGE_G2P)
    _, vel, c = o.get_particles()
    expect = V0 + pos @ G.T
    assert np.allclose(vel[:, :2], expect[:, :2], rtol=0, atol=1e-12)
    assert np.all(vel[:, 2] == 0.0)
    if transfer == capi.APIC:
        assert np.allclose(c[:, :6], np.tile(G[:2].reshape(-1), (len(pos), 1)), rtol=0, atol=1e-10)
        assert np.all(c[:, 6:] == 7.0)


def test_2d_apic_p2g_reproduces_affine_field_at_anisotropic_faces():
    """v_p + C (x_f - x_p) with v_p = v0 + G x_p, C = G gives v0 + G x_f at every touched face: the
    z face positions are (z + 0.5) and (z + 1) times cellD.z."""
    grid = scenes.grid_2d(12)
    o = mk(grid, transfer=capi.APIC)
    pos = rand_2d(grid, 4000, 2, z_any=True)
    o.set_particles(pos, V0 + pos @ G.T, np.tile(G.reshape(-1), (len(pos), 1)))
    o.run_stage(capi.STAGE_P2G)
    fp = face_positions(grid)
    for a, f in enumerate((capi.FIELD_U, capi.FIELD_V, capi.FIELD_W)):
        v = o.get_grid(f)
        nz = v != 0
        assert nz.sum() > 200
        assert np.allclose(v[nz], V0[a] + fp[a][nz] @ G[a], rtol=0, atol=1e-12)


def test_2d_bin_keys_and_water_cells():
    """Bin keys use floor(z / cellD.z); with z = depth / 2 every particle's cell is in the middle layer,
    and MARK makes exactly the cells that hold particles WATER."""
    grid = scenes.grid_2d(10)
    o = mk(grid)
    pos = rand_2d(grid, 3000, 6, z_any=True)
    o.set_particles(pos)
    count, start, perm = o.bin()
    cell = np.floor(pos / np.array([grid.cell_d, grid.cell_d, grid.cell_dz])).astype(int)
    assert np.all(cell[:, 2] == 1)
    key = (cell[:, 0] * 10 + cell[:, 1]) * 3 + cell[:, 2]
    assert np.array_equal(count, np.bincount(key, minlength=grid.nc))
    assert np.array_equal(perm, np.argsort(key, kind="stable"))
    o.run_stage(capi.STAGE_P2G)
    o.run_stage(capi.STAGE_MARK)
    lab = o.get_grid(capi.FIELD_LABEL)
    assert np.array_equal(lab == capi.WATER, count > 0)


def test_2d_wall_bounce_in_z_ignores_the_radius():
    """The 2D clamp box in z is [cellD.z, depth - cellD.z] (gridLow.z without r, simulator.cpp:140):
    a particle flying at +z reverses at that wall with restitution 0.3 (the 3D closed form)."""
    grid = scenes.grid_2d(16)
    o = mk(grid)
    hi = DEPTH - grid.cell_dz
    z0, vz, dt = hi - 0.004, 2.0, 0.01
    o.set_particles(np.array([[0.5, 0.5, z0]]), np.array([[0.0, 0.0, vz]]))
    o.run_stage(capi.STAGE_ADVECT, dt)
    p, v, _ = o.get_particles()
    t_hit = (hi - z0) / vz
    expect = z0 + vz * t_hit * 0.999 + (-0.3 * vz) * (dt - 0.999 * t_hit)
    assert p[0, 2] == pytest.approx(expect, abs=1e-15)
    assert v[0, 2] == pytest.approx(-0.3 * vz, abs=1e-15)


@pytest.mark.parametrize("r_cells", [0.25, 0.8])
def test_2d_push_apart_all_pairs_then_fixed_z(r_cells):
    """The binned pass in 2D finds every pair closer than D with the per-axis reach (D spans several
    z cells here) and applies the offsets of the all-pairs sum; x, y are clamped and z is set to
    depth / 2 (hashedParticles.cpp:96-101)."""
    n = 16
    grid = capi.Grid(n, n, 3, 1.0 / n, r_cells / n, two_d=1, cell_d_z=DEPTH / 3)
    rng = np.random.default_rng(8)
    pos = rng.uniform(0.3, 0.45, (1200, 3))
    pos[:, 2] = rng.uniform(grid.cell_dz, 2 * grid.cell_dz, len(pos))
    pos[:20, :2] = rng.uniform(0.0, 1.0, (20, 2)) * 0.05 + grid.cell_d
    o = oracle.Oracle(grid, scenes.params(push_apart_enabled=1))
    o.set_particles(pos, np.zeros_like(pos))
    o.run_stage(capi.STAGE_PUSH_APART)
    got, _, _ = o.get_particles()
    D = 2 * grid.particle_r
    d = pos[:, None, :] - pos[None, :, :]
    d2 = (d ** 2).sum(-1)
    near = (d2 <= D * D) & (d2 >= 1e-8)
    np.fill_diagonal(near, False)
    dist = np.sqrt(np.where(near, d2, 1.0))
    off = np.where(near[..., None], d * ((D - dist) / dist)[..., None] * 0.5, 0.0).sum(1)
    lo = grid.cell_d + 1.01 * grid.particle_r
    expect = np.clip(pos + off, lo, 1.0 - lo)
    expect[:, 2] = DEPTH / 2
    assert near.sum() > 1000
    np.testing.assert_allclose(got, expect, rtol=0, atol=1e-13)


def test_2d_sphere_is_a_disc_of_the_middle_layer():
    """A sphere's rasterisation range in z is clamped to [1, nz - 2] = the middle layer, its centre z is
    pos.z / cellD.z and its radius counts x cells (macGrid.cpp:235-240): a sphere at z = depth / 2
    solidifies the disc (x + .5 - cx)^2 + (y + .5 - cy)^2 < r^2 of layer 1 and nothing else."""
    n = 20
    grid = scenes.grid_2d(n)
    sphere = capi.make_obstacle(capi.OBS_SPHERE, (0.45, 0.55, DEPTH / 2), r=0.23)
    o = mk(grid)
    o.set_particles(rand_2d(grid, 10, 1)[:0])
    o.set_obstacles([sphere])
    o.run_stage(capi.STAGE_OBSTACLES, DT)
    lab = o.get_grid(capi.FIELD_LABEL).reshape(n, n, 3)
    X, Y = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    disc = (X + 0.5 - 0.45 * n) ** 2 + (Y + 0.5 - 0.55 * n) ** 2 < (0.23 * n) ** 2
    disc &= (X >= 1) & (Y >= 1) & (X <= n - 2) & (Y <= n - 2)
    assert disc.sum() > 50
    assert np.array_equal(lab[:, :, 1] == capi.SOLID, disc)
    assert not np.any(lab[:, :, 0] == capi.SOLID) and not np.any(lab[:, :, 2] == capi.SOLID)


def test_2d_box_uses_cell_d_z_for_its_depth():
    """A box's z extent is size.z / cellD.z cells around pos.z / cellD.z (rounded, clamped to
    [1, nz - 1], macGrid.cpp:266-269): a box as deep as the domain covers layer 1 only; a box depth / 3
    deep centred on layer 2 spans [2, 2), no layer (counted in x cells it would cover layer 1); a box
    as deep as the domain centred on layer 0 spans [1, 2), layer 1 (in x cells: [1, 1), none)."""
    n = 16
    grid = scenes.grid_2d(n)
    o = mk(grid)
    o.set_particles(np.zeros((0, 3)))
    o.set_obstacles([capi.make_obstacle(capi.OBS_BOX, (0.5, 0.5, DEPTH / 2), size=(0.25, 0.25, DEPTH))])
    o.run_stage(capi.STAGE_OBSTACLES, DT)
    lab = o.get_grid(capi.FIELD_LABEL).reshape(n, n, 3)
    assert np.array_equal(np.argwhere(lab == capi.SOLID)[:, 2], np.ones((lab == capi.SOLID).sum(), int))
    assert (lab[:, :, 1] == capi.SOLID).sum() == 4 * 4
    o2 = mk(grid)
    o2.set_particles(np.zeros((0, 3)))
    o2.set_obstacles([capi.make_obstacle(capi.OBS_BOX, (0.5, 0.5, DEPTH * 5 / 6), size=(0.25, 0.25, DEPTH / 3))])
    o2.run_stage(capi.STAGE_OBSTACLES, DT)
    assert not np.any(o2.get_grid(capi.FIELD_LABEL) == capi.SOLID)
    o3 = mk(grid)
    o3.set_particles(np.zeros((0, 3)))
    o3.set_obstacles([capi.make_obstacle(capi.OBS_BOX, (0.5, 0.5, DEPTH / 6), size=(0.25, 0.25, DEPTH))])
    o3.run_stage(capi.STAGE_OBSTACLES, DT)
    lab3 = o3.get_grid(capi.FIELD_LABEL).reshape(n, n, 3)
    assert np.array_equal(lab3 == capi.SOLID, lab == capi.SOLID)


def _dam_2d(n=16, **kw):
    sc = scenes.dam_break_2d(n, capi.FLIP, sphere=True)
    o = oracle.Oracle(sc.grid, scenes.params(push_apart_enabled=1, **kw))
    o.set_particles(sc.particles().astype(np.float64))
    o.set_obstacles(sc.obstacles)
    return sc.grid, o


def test_2d_pcg_is_the_5_point_solve_of_the_middle_layer():
    """The z layers 0 and 2 are SOLID walls, so the 7-point matrix of the WATER cells is the 5-point
    Laplacian of layer 1 (diagonal = non-SOLID in-plane neighbours); the PCG pressure equals its
    direct solve and the projected field is divergence-free in the plane."""
    grid, o = _dam_2d(24, tol=1e-12)
    for _ in range(3):
        o.step(DT)
    for st in (capi.STAGE_P2G, capi.STAGE_MARK, capi.STAGE_OBSTACLES, capi.STAGE_BORDERS, capi.STAGE_POST_P2G):
        o.run_stage(st, DT)
    n = grid.nx
    lab = o.get_grid(capi.FIELD_LABEL).reshape(n, n, 3)
    assert np.all(lab[:, :, 0] == capi.SOLID) and np.all(lab[:, :, 2] == capi.SOLID)
    o.run_stage(capi.STAGE_SOLVE, DT)
    A, cells = laplacian(lab, DT / grid.cell_d ** 2)
    assert len(cells) > 60 and np.all(cells[:, 2] == 1)
    assert A.diagonal().max() <= 4 * DT / grid.cell_d ** 2
    b = o.get_grid(capi.FIELD_RHS).reshape(n, n, 3)[tuple(cells.T)]
    p = o.get_grid(capi.FIELD_PRESSURE).reshape(n, n, 3)[tuple(cells.T)]
    p_direct = spla.spsolve(A.tocsc(), b)
    assert np.abs(p - p_direct).max() <= 1e-9 * np.abs(p_direct).max()
    u, v = (o.get_grid(f).reshape(n, n, 3)[:, :, 1] for f in (capi.FIELD_U2, capi.FIELD_V2))
    div = np.zeros((n, n))
    div[1:-1, 1:-1] = (u[1:-1, 1:-1] - u[:-2, 1:-1] + v[1:-1, 1:-1] - v[1:-1, :-2]) / grid.cell_d
    assert np.abs(div[cells[:, 0], cells[:, 1]]).max() <= 1e-8 * np.abs(b).max() * grid.cell_d / DT


def test_2d_steps_keep_the_particle_plane():
    """Whole 2D steps with push-apart: every particle ends each step at z = depth / 2 with v.z = 0,
    inside the x-y clamp box, and the sphere's layer-1 disc holds none of them."""
    grid, o = _dam_2d()
    lo = grid.cell_d + 1.01 * grid.particle_r
    n0 = o.num_particles()
    for _ in range(6):
        o.step(DT)
        pos, vel, _ = o.get_particles()
        assert len(pos) == n0
        assert np.all(pos[:, 2] == DEPTH / 2) and np.all(vel[:, 2] == 0.0)
        assert pos[:, :2].min() >= lo and pos[:, :2].max() <= 1 - lo
        assert np.isfinite(vel).all()
    assert np.abs(vel[:, 1]).max() > 0.05   # the block is falling
