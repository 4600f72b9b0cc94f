"""Training data and training loop for the learned pressure preconditioner.

The preconditioner is a multigrid V-cycle whose coarsest level (at most 16
cells a side) is solved by the network (include/fsim.h, FSIM_PRECOND_LEARNED).
Its training samples are that level's problems as the GPU solver meets them:
seeded random dam-break and sphere-drop scenes are stepped on the GPU with
MIC(0)-PCG, and at chosen steps the pressure solve is run once more by
fsim_learned_harvest (flexible PCG with the V-cycle and an accurate
coarsest-level solve), which returns each coarsest-level right-hand side b,
its solution x and the level's operator (Galerkin coefficients on the
network's scale).  The network is fully convolutional: it is trained on a
context whose grid is the coarsest level's and its parameters are then set
on a context of any size with Simulation.learned_set_params.
"""
from __future__ import annotations

import numpy as np

from . import capi, scenes

DT = 1.0 / 240.0
LOSS_MSE, LOSS_ENERGY = 0, 1
PREP_STAGES = (capi.STAGE_ADVECT, capi.STAGE_PUSH_OUT, capi.STAGE_P2G, capi.STAGE_MARK, capi.STAGE_OBSTACLES,
               capi.STAGE_BORDERS, capi.STAGE_POST_P2G)

# (input channels, output channels, level) of the 3x3x3 layers, in order
# (fsim_learned.hip kLayers); the 1x1 output layer maps 16 channels to 1.
LAYERS = [(8, 16, 0), (16, 32, 1), (32, 64, 2), (64, 128, 3), (192, 64, 2), (96, 32, 1), (48, 16, 0)]


def flops_per_apply(shape) -> float:
    """Multiply-add FLOPs (2 per MAC) of one network evaluation on a grid of `shape` cells, as executed:
    every dimension padded to a multiple of 8, all 8 input channels counted."""
    d0 = [(n + 7) // 8 * 8 for n in shape]
    total = 0.0
    for cin, cout, lev in LAYERS:
        vox = float(np.prod([n >> lev for n in d0]))
        total += 2.0 * 27 * cin * cout * vox
    return total + 2.0 * 16 * float(np.prod(d0))


def random_scene(rng: np.random.Generator, n: int, seed: int) -> scenes.Scene:
    """A seeded random dam break (one block of random size, or two opposing ones) or sphere drop (a ball
    of water above a pool)."""
    kind = int(rng.integers(0, 3))
    if kind == 0:
        blocks = [[0.0, 0.0, 0.0, rng.uniform(0.2, 0.6), rng.uniform(0.3, 0.8), 1.0]]
    elif kind == 1:
        w, h = rng.uniform(0.2, 0.45), rng.uniform(0.3, 0.7)
        blocks = [[0.0, 0.0, 0.0, w, h, 1.0], [1.0 - w, 0.0, 0.0, 1.0, h, 1.0]]
    else:
        r = rng.uniform(0.08, 0.18)
        c = (rng.uniform(0.3, 0.7), rng.uniform(0.55, 0.75), rng.uniform(0.3, 0.7))
        blocks = [[0.0, 0.0, 0.0, 1.0, rng.uniform(0.15, 0.35), 1.0],
                  [c[0] - r, c[1] - r, c[2] - r, c[0] + r, c[1] + r, c[2] + r]]
    return scenes.Scene(f"random_{kind}_{n}", n, capi.FLIP, blocks, [], seed)


def harvest_samples(n: int, scenes_count: int, harvests_per_scene: int, per_solve: int = 12, seed: int = 0,
                    max_steps: int = 30, log=None):
    """Coarsest-level samples (B [k, Nc'], OP [k, 4, Nc'], X [k, Nc'], dims) from scenes_count random scenes
    on an n^3 grid, harvests_per_scene solves each (at random steps up to max_steps), at most per_solve
    samples per solve."""
    rng = np.random.default_rng(seed)
    grid = scenes.grid(n)
    B, OP, X = [], [], []
    dims = None
    for s in range(scenes_count):
        sc = random_scene(rng, n, seed * 1000 + s)
        pos = capi.scene_dam_break(grid, sc.blocks, seed=sc.seed)
        if len(pos) == 0:
            continue
        sim = capi.Simulation(grid, scenes.params())
        sim.set_particles(pos)
        at = sorted(rng.choice(np.arange(1, max_steps), size=harvests_per_scene, replace=False))
        step = 0
        for a in at:
            while step < a - 1:
                sim.step(DT)
                step += 1
            for st in PREP_STAGES:
                sim.run_stage(st, DT)
            b, op, x, dims = sim.learned_harvest(DT, per_solve)
            for k in range(len(b)):
                B.append(b[k])
                OP.append(op)
                X.append(x[k])
            # finish the step the harvest solved (its pressure is applied)
            sim.run_stage(capi.STAGE_EXTRAPOLATE, DT)
            sim.run_stage(capi.STAGE_G2P, DT)
            step += 1
        sim.close()
        if log:
            log(f"scene {s}: {len(B)} samples")
    return np.stack(B), np.stack(OP), np.stack(X), tuple(dims)


def train(dims, B, OP, X, epochs: int, batch: int = 8, lr: float = 1e-3, seed: int = 0,
          params: np.ndarray | None = None, log=None, loss: int = LOSS_MSE) -> tuple[np.ndarray, list[float]]:
    """Adam over shuffled mini-batches on a context whose grid is the coarsest level's (dims); returns the
    parameters and the per-epoch mean loss.  loss: LOSS_MSE (masked MSE against x / s) or LOSS_ENERGY
    (e^T L e / t^T L t with the sample's operator)."""
    rng = np.random.default_rng(seed)
    grid = capi.Grid(*dims, 1.0 / dims[0], 0.25 / dims[0])
    sim = capi.Simulation(grid, scenes.params())
    sim.set_option(capi.OPT_LEARNED_LOSS, loss)
    if params is None:
        sim.learned_init(seed)
    else:
        sim.learned_set_params(params)
    history = []
    k = len(B)
    for ep in range(epochs):
        order = rng.permutation(k)
        losses = []
        for i in range(0, k, batch):
            idx = order[i:i + batch]
            losses.append(sim.learned_train_batch_op(B[idx], OP[idx], X[idx].astype(np.float64), lr))
        history.append(float(np.mean(losses)))
        if log:
            log(f"epoch {ep}: loss {history[-1]:.4e}")
    out = sim.learned_get_params()
    sim.close()
    return out, history


def save(path: str, params: np.ndarray) -> None:
    np.save(path, np.asarray(params, np.float32), allow_pickle=False)


def load(path: str) -> np.ndarray:
    p = np.load(path, allow_pickle=False)
    if p.size != capi.learned_num_params():
        raise ValueError(f"{path}: {p.size} parameters, expected {capi.learned_num_params()}")
    return p.astype(np.float32)


def forward_timing(n: int, reps: int = 10, seed: int = 1, conv: int = 1) -> dict:
    """Wall time of fsim_learned_forward on an n^3 grid (random right-hand side, every cell WATER,
    initial parameters), host copies included; for profiling the network's kernels.
    conv: FSIM_OPT_LEARNED_CONV (1 LDS-staged convolutions where they fit, 0 cached loads only)."""
    import time
    grid = scenes.grid(n)
    sim = capi.Simulation(grid, scenes.params())
    sim.set_option(capi.OPT_LEARNED_CONV, conv)
    sim.learned_init(seed)
    b = np.random.default_rng(seed).standard_normal(grid.nc).astype(np.float32)
    lab = np.full(grid.nc, capi.WATER, np.uint8)
    sim.learned_forward(b, lab)
    t0 = time.perf_counter()
    for _ in range(reps):
        out = sim.learned_forward(b, lab)
    dt = (time.perf_counter() - t0) / reps
    sim.close()
    return {"grid": [n] * 3, "reps": reps, "conv": conv, "ms_per_forward_with_copies": dt * 1e3,
            "gflop_per_forward": flops_per_apply(grid.shape) / 1e9, "checksum": float(np.abs(out).sum())}


def solve_timing(n: int, coarse: int, reps: int = 3, seed: int = 1, graph: int = 1) -> dict:
    """Pressure solves of one state of the n^3 FLIP double dam break (3 MIC(0)-PCG steps in), repeated
    `reps` times with the learned multigrid FPCG and the given coarsest-level mode (capi.MG_COARSE_*;
    NETWORK uses initial parameters from `seed`, so its iteration count is not a trained one): for
    profiling the V-cycle's kernels; graph: FSIM_OPT_MG_GRAPH.  Returns the solves' ms and iteration counts."""
    import time
    sc = scenes.double_dam_break(n, capi.FLIP, seed=1)
    sim = capi.Simulation(sc.grid, scenes.params(transfer=sc.transfer))
    sim.set_particles(sc.particles())
    for _ in range(3):
        st = sim.step(DT)
    state = sim.get_particles()
    k_mic = st.pcg_iterations
    sim.set_params(scenes.params(transfer=sc.transfer, precond=capi.PRECOND_LEARNED))
    sim.set_option(capi.OPT_MG_COARSE, coarse)
    sim.set_option(capi.OPT_LEARNED_REFRESH, 0)
    sim.set_option(capi.OPT_MG_GRAPH, graph)
    sim.learned_init(seed)
    out = []
    t0 = time.perf_counter()
    for _ in range(reps):
        sim.set_particles(*state)
        sim.set_option(capi.OPT_LEARNED_MIC_COUNT, 10 * k_mic)
        st = sim.step(DT)
        out.append({"solve_ms": round(st.solve_ms, 3), "iterations": st.learned_iterations, "stop": st.learned_stop})
    wall = time.perf_counter() - t0
    sim.close()
    return {"grid": [n] * 3, "coarse": coarse, "graph": graph, "mic0_iterations": k_mic, "solves": out,
            "wall_s": round(wall, 2)}


if __name__ == "__main__":
    import argparse
    import json

    ap = argparse.ArgumentParser(description="Time the learned network's forward pass (fsim_learned_forward), "
                                             "or (--solve) learned multigrid pressure solves")
    ap.add_argument("--size", type=int, default=128)
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--conv", type=int, default=1, choices=(0, 1))
    ap.add_argument("--solve", action="store_true", help="time solves of a double dam break state instead")
    ap.add_argument("--coarse", type=int, default=capi.MG_COARSE_SMOOTH,
                    help="with --solve: coarsest-level mode (0 network, 1 smooth, 2 accurate)")
    ap.add_argument("--no-graph", action="store_true", help="with --solve: launch the V-cycle's kernels one by one")
    a = ap.parse_args()
    if a.solve:
        print(json.dumps(solve_timing(a.size, a.coarse, a.reps, graph=0 if a.no_graph else 1)), flush=True)
    else:
        print(json.dumps(forward_timing(a.size, a.reps, conv=a.conv)), flush=True)
