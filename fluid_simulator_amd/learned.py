"""Training data and training loop for the learned pressure preconditioner.

Samples are (b, label, p) triples of real pressure solves: seeded random
dam-break scenes (1-3 boxes) are stepped on the GPU with MIC(0)-PCG and each
step's right-hand side, cell labels and converged pressure are recorded.  The
network (include/fsim.h, fsim_learned_*) is fully convolutional, so it is
trained on a small grid and its parameters are then set on a context of any
size with Simulation.learned_set_params.
"""
from __future__ import annotations

import numpy as np

from . import capi, scenes

DT = 1.0 / 240.0
LOSS_MSE, LOSS_ENERGY = 0, 1

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


def random_blocks(rng: np.random.Generator, n_max: int = 3):
    blocks = []
    for _ in range(int(rng.integers(1, n_max + 1))):
        lo = rng.uniform(0.0, 0.6, 3)
        hi = np.minimum(lo + rng.uniform(0.2, 0.6, 3), 1.0)
        lo[1] = 0.0   # boxes stand on the floor
        blocks.append([*lo, *hi])
    return blocks


def generate_samples(n: int, scenes_count: int, steps_per_scene: int, seed: int = 0, residuals_per_solve: int = 0):
    """Arrays B [k, Nc] float32, L [k, Nc] uint8, P [k, Nc] float64 from the solves on an n^3 grid,
    each followed by residuals_per_solve residual_samples pairs."""
    rng = np.random.default_rng(seed)
    grid = scenes.grid(n)
    B, L, P = [], [], []
    for s in range(scenes_count):
        pos = capi.scene_dam_break(grid, random_blocks(rng), seed=seed * 1000 + s)
        if len(pos) == 0:
            continue
        sim = capi.Simulation(grid, scenes.params())
        sim.set_particles(pos)
        for _ in range(steps_per_scene):
            st = sim.step(DT)
            if st.early_exit:
                continue
            b, lab, p = sim.get_grid(capi.FIELD_RHS), sim.get_grid(capi.FIELD_LABEL), sim.get_grid(capi.FIELD_PRESSURE)
            B.append(b)
            L.append(lab)
            P.append(p)
            if residuals_per_solve:
                rb, rl, rp = residual_samples(sim, lab, p, residuals_per_solve, seed=seed * 7919 + len(B))
                B.extend(rb)
                L.extend(rl)
                P.extend(rp)
        sim.close()
    return grid, np.stack(B), np.stack(L), np.stack(P)


def _smooth(f: np.ndarray, k: int) -> np.ndarray:
    for _ in range(k):
        f = (f + np.roll(f, 1, 0) + np.roll(f, -1, 0) + np.roll(f, 1, 1) + np.roll(f, -1, 1) + np.roll(f, 1, 2)
             + np.roll(f, -1, 2)) / 7.0
    return f


def residual_samples(sim: capi.Simulation, label: np.ndarray, p: np.ndarray, count: int, seed: int = 0):
    """Pairs (A e, e) for error fields e on the WATER cells of the simulation's last solve:
    white noise smoothed 0-11 times, every third one mixed with the solve's pressure p.
    Inside flexible PCG the network sees residuals, not only right-hand sides; A e comes from
    the library's own SpMV (fsim_solver_apply), so every pair is exact."""
    rng = np.random.default_rng(seed)
    shape = sim.grid.shape
    fl = (label == capi.WATER)
    pmax = max(float(np.abs(p).max()), 1e-30)
    B, P = [], []
    for k in range(count):
        e = _smooth((rng.standard_normal(sim.grid.nc) * fl).reshape(shape), int(rng.integers(0, 12))).reshape(-1) * fl
        if k % 3 == 0:
            e = e + rng.uniform(0.2, 1.0) * p / pmax * np.abs(e).max()
        B.append(sim.solver_apply(capi.OP_APPLY_A, DT, e).astype(np.float32))
        P.append(e)
    return np.stack(B), np.stack([label] * count), np.stack(P)


def train(grid: capi.Grid, B, L, P, epochs: int, batch: int = 4, lr: float = 1e-3, seed: int = 0,
          params: np.ndarray | None = None, log=None, loss: int = LOSS_MSE) -> tuple[np.ndarray, list[float]]:
    """Adam over shuffled mini-batches; returns the parameters and the per-epoch mean loss.
    loss: LOSS_MSE (masked MSE against p/s) or LOSS_ENERGY (e^T A e / t^T A t, the norm CG minimises)."""
    rng = np.random.default_rng(seed)
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
            losses.append(sim.learned_train_batch(B[idx], L[idx], P[idx], lr))
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


if __name__ == "__main__":
    import argparse
    import json

    ap = argparse.ArgumentParser(description="Time the learned network's forward pass (fsim_learned_forward)")
    ap.add_argument("--size", type=int, default=128)
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--conv", type=int, default=1, choices=(0, 1))
    a = ap.parse_args()
    print(json.dumps(forward_timing(a.size, a.reps, conv=a.conv)), flush=True)
