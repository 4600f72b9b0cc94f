"""Shared helpers for the parity tests (GPU path vs the FP64 oracle)."""
from __future__ import annotations

import numpy as np

from fluid_simulator_amd import capi, scenes


def vector_rel_err(sim, ref, fields) -> float:
    """max over the components of |gpu - ref|, divided by the largest |ref| of all of them: the error of a
    vector field (u, v, w faces) relative to that field's scale.  A component that vanishes in exact
    arithmetic (w of a scene uniform in z) then carries no scale of its own."""
    g = [np.asarray(sim.get_grid(f), np.float64) for f in fields]
    r = [np.asarray(ref.get_grid(f), np.float64) for f in fields]
    scale = max(max(float(np.abs(a).max()) for a in r), 1e-30)
    return max(float(np.abs(a - b).max()) for a, b in zip(g, r)) / scale


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


F_FACES = (capi.FIELD_U, capi.FIELD_V, capi.FIELD_W)
F_FACES2 = (capi.FIELD_U2, capi.FIELD_V2, capi.FIELD_W2)


def copy_grid_to_oracle(sim: capi.Simulation, ref, fields) -> None:
    """The GPU's FP32 grid fields, exactly, into the oracle."""
    for f in fields:
        ref.set_grid(f, sim.get_grid(f).astype(np.float64))


def staged_step(sim: capi.Simulation, ref, dt: float, spawn: bool = False, stop: bool = False,
                push_apart: bool = False) -> dict:
    """One step of the reference's stage order (simulator.cpp:51-100) on both
    backends, stage by stage, each stage started from identical inputs: after
    every stage the oracle takes the GPU's FP32 result (particles or grid
    fields).  So each stage's error is its own, not carried over from the
    FP32 rounding of the stages before it.  Returns, per stage, the largest
    relative error of its outputs (max |gpu - oracle| / max |oracle|), the
    label and binning mismatches (counts of differing entries) and the
    solver statistics; the GPU state is that of its own step."""
    e = {}
    sync_oracle_from_gpu(sim, ref)
    if spawn:
        sim.run_stage(capi.STAGE_SPAWN, dt)
        ref.run_stage(capi.STAGE_SPAWN, dt)
        e["spawn_count_diff"] = abs(sim.num_particles() - ref.num_particles())
        p_g, v_g, _ = sim.get_particles()
        p_r, v_r, _ = ref.get_particles()
        e["spawn_pos_abs"] = float(np.abs(p_g - p_r).max()) if len(p_g) else 0.0
        sync_oracle_from_gpu(sim, ref)
    stages = [(capi.STAGE_ADVECT, "advect")]
    if push_apart:   # pushParticlesApart runs between advection and push-out (simulator.cpp:61-64)
        stages.append((capi.STAGE_PUSH_APART, "push_apart"))
    stages.append((capi.STAGE_PUSH_OUT, "push_out"))
    for stage, name in stages:
        sim.run_stage(stage, dt)
        ref.run_stage(stage, dt)
        sim.run_stage(capi.STAGE_BIN, dt)   # drops the particles a sink removed, as the oracle's advection does
        p_g, v_g, _ = sim.get_particles()
        p_r, v_r, _ = ref.get_particles()
        e[f"{name}_count_diff"] = abs(len(p_g) - len(p_r))
        if len(p_g) == len(p_r) and len(p_g):
            e[f"{name}_pos_abs"] = float(np.abs(p_g - p_r).max())   # domain [0, 1]: absolute = relative
            e[f"{name}_vel"] = field_rel_err(v_g, v_r)
        sync_oracle_from_gpu(sim, ref)
    if stop:
        sim.run_stage(capi.STAGE_STOP, dt)
        ref.run_stage(capi.STAGE_STOP, dt)
    # binning + P2G (+ density) + marking from the same FP32 positions
    sim.run_stage(capi.STAGE_P2G, dt)
    ref.run_stage(capi.STAGE_P2G, dt)
    cnt_r, start_r, perm_r = ref.bin()
    e["bin_count_mismatch"] = int(np.count_nonzero(sim.get_grid(capi.FIELD_CELL_COUNT) != cnt_r))
    e["bin_start_mismatch"] = int(np.count_nonzero(sim.get_grid(capi.FIELD_CELL_START) != start_r))
    e["bin_perm_mismatch"] = int(np.count_nonzero(sim.slot_ids() != perm_r))
    e["p2g"] = vector_rel_err(sim, ref, F_FACES)
    sim.run_stage(capi.STAGE_MARK, dt)
    ref.run_stage(capi.STAGE_MARK, dt)
    # the oracle splats the density while marking (markFluidCellsAndCalculateParticleDensities,
    # simulator.cpp:354-368); the GPU fuses it into P2G, so both hold it from here on
    e["density"] = field_rel_err(sim.get_grid(capi.FIELD_DENSITY), ref.get_grid(capi.FIELD_DENSITY))
    e["mark_label_mismatch"] = int(np.count_nonzero(sim.get_grid(capi.FIELD_LABEL) != ref.get_grid(capi.FIELD_LABEL)))
    copy_grid_to_oracle(sim, ref, F_FACES + (capi.FIELD_DENSITY,))
    # obstacles, borders, v2 + gravity
    for stage in (capi.STAGE_OBSTACLES, capi.STAGE_BORDERS, capi.STAGE_POST_P2G):
        sim.run_stage(stage, dt)
        ref.run_stage(stage, dt)
    e["prep_label_mismatch"] = int(np.count_nonzero(sim.get_grid(capi.FIELD_LABEL) != ref.get_grid(capi.FIELD_LABEL)))
    e["prep"] = max(vector_rel_err(sim, ref, F_FACES), vector_rel_err(sim, ref, F_FACES2))
    copy_grid_to_oracle(sim, ref, F_FACES + F_FACES2)
    # pressure solve from the same faces
    sim.run_stage(capi.STAGE_SOLVE, dt)
    ref.run_stage(capi.STAGE_SOLVE, dt)
    e["rhs"] = field_rel_err(sim.get_grid(capi.FIELD_RHS), ref.get_grid(capi.FIELD_RHS))
    e["pressure"] = field_rel_err(sim.get_grid(capi.FIELD_PRESSURE), ref.get_grid(capi.FIELD_PRESSURE))
    e["projected"] = vector_rel_err(sim, ref, F_FACES2)
    st_g, st_r = sim.stats(), ref.stats()
    e["pcg_gpu"], e["pcg_oracle"] = st_g.pcg_iterations, st_r.pcg_iterations
    e["residual_gpu"], e["residual_oracle"] = st_g.residual_rel_l2, st_r.residual_rel_l2
    e["b_norm"] = st_g.b_norm_l2
    copy_grid_to_oracle(sim, ref, F_FACES2)
    sim.run_stage(capi.STAGE_EXTRAPOLATE, dt)
    ref.run_stage(capi.STAGE_EXTRAPOLATE, dt)
    e["extrapolate"] = vector_rel_err(sim, ref, F_FACES2)
    copy_grid_to_oracle(sim, ref, F_FACES2)
    sim.run_stage(capi.STAGE_G2P, dt)
    ref.run_stage(capi.STAGE_G2P, dt)
    p_g, v_g, c_g = sim.get_particles()
    p_r, v_r, c_r = ref.get_particles()
    e["g2p_vel"] = field_rel_err(v_g, v_r)
    # 2D: the particles sit on the cell-centre plane z = depth / 2, the peak of the z tent, where
    # trilinearInterpollGradient's sign(p.z - face.z) (interpolation.h:23-35) decides the z component
    # of the u and v gradients.  For the depth 0.1 of these scenes the reference (FP64 positions) and
    # the GPU (FP32 positions and face positions) both see p.z - face.z = 0 there; the oracle, fed the
    # FP32 z = 0.05f against its FP64 face z = 0.05, sees +7e-10 and takes the other side.  Those two components (c[0].z, c[1].z)
    # have no effect on P2G (the faces of the particle's layer are at its z, the other layers weigh
    # 0), so they are left out of the 2D comparison; c[2] (never written in 2D) is compared.
    cols = [0, 1, 3, 4, 6, 7, 8] if sim.grid.two_d else list(range(9))
    e["g2p_c"] = field_rel_err(c_g[:, cols], c_r[:, cols]) if np.abs(c_r[:, cols]).max() > 0 else 0.0
    return e


STAGE_FIELD_KEYS = ("p2g", "density", "prep", "rhs", "pressure", "projected", "extrapolate", "g2p_vel")
STAGE_EXACT_KEYS = ("bin_count_mismatch", "bin_start_mismatch", "bin_perm_mismatch", "mark_label_mismatch",
                    "prep_label_mismatch", "advect_count_diff", "push_apart_count_diff", "push_out_count_diff")
