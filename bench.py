#!/usr/bin/env python3
"""Benchmark of the Simulator step on one MI355X.

Workload: the 256^3 FLIP (0.95) double dam break (two blocks x<0.4 and x>0.6,
y<0.6, 8 particles per fluid cell, seed 1), dt = 1/240, MIC(0)-PCG to
||r||_2 <= 1e-6 ||b||_2.  A "step" is one Simulator::simulate(dt): advection,
binning, P2G, labels, pressure solve, extrapolation, G2P, with the particle
and grid state resident in HBM.

Prints one JSON line: steps/s (value), plus the roofline of the dominant
kernel measured with HIP events on the library's stream, the FP64 CPU
oracle timed on a bounded sample of the same workload, a 20-step 128^3
GPU-vs-oracle error trace, the second metric, pressure-solve ms to 1e-6 with
MIC(0)-PCG vs the learned multigrid FPCG on the same right-hand side at 128^3
and 256^3 (network trained in this run, on the GPU), and the 256^3 scene with
solid boxes and a particle source run with both solvers.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from fluid_simulator_amd import capi, learned, scenes  # noqa: E402

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E

# Algorithmic bytes per unit for each kernel group (FP32 fields, uint8 labels,
# each array counted once per launch, no cache credit).  Units: f = WATER
# cells, c = all cells, p = particles.
ALGO_BYTES = {
    "sweep_fwd": ("f", 13.0),      # R r 4, R P 4, R label 1, W q 4
    "sweep_bwd": ("f", 13.0),      # R q 4, R P 4, R label 1, W z 4
    "mic_factor": ("f", 5.0),      # R label 1, W P 4
    "spmv": ("f", 17.0),           # direction step: R label 1, R z 4, R s 4, W s' 4, W q 4
    "update": ("f", 13.0),         # R label 1, R q 4, R/W r 8
    "dot": ("f", 8.0),             # R z 4, R r 4
    "p2g": ("p+c", (24.0, 12.0)),  # R particle pos+vel 24 B/p, W u,v,w 12 B/cell
    "g2p": ("p+c", (36.0, 24.0)),  # R pos,vel W vel 36 B/p, R u,v,w,u2,v2,w2 24 B/cell
    "move": ("p", 36.0),           # R pos,vel W pos (vel in place) 36 B/p
    "extrapolate": ("c", 54.0),
}


def algo_bytes(name: str, nf: float, nc: int, np_: int) -> float | None:
    if name not in ALGO_BYTES:
        return None
    unit, b = ALGO_BYTES[name]
    if unit == "f":
        return b * nf
    if unit == "c":
        return b * nc
    if unit == "p":
        return b * np_
    return b[0] * np_ + b[1] * nc


def divergence_l2(sim, grid) -> float:
    """L2 norm over WATER cells of the MAC divergence of the projected velocity (u2, v2, w2); a cell's
    faces[k] is its +k face, so div(i) = (u2(i) - u2(i-x) + v2(i) - v2(i-y) + w2(i) - w2(i-z)) / dx."""
    shape = grid.shape
    lab = sim.get_grid(capi.FIELD_LABEL).reshape(shape)
    d = np.zeros(shape, np.float64)
    for ax, f in enumerate((capi.FIELD_U2, capi.FIELD_V2, capi.FIELD_W2)):
        v = sim.get_grid(f).reshape(shape).astype(np.float64)
        d += v - np.roll(v, 1, axis=ax)
    d /= grid.cell_d
    return float(np.sqrt(np.sum(np.where(lab == capi.WATER, d, 0.0) ** 2)))


def host_info() -> dict:
    """nproc, the CPUs this process may run on, and the CPU model (lscpu, else /proc/cpuinfo)."""
    import subprocess
    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        model = next((ln.split(":", 1)[1].strip() for ln in out.splitlines() if ln.startswith("Model name")), None)
    except (OSError, subprocess.SubprocessError):
        pass
    if model is None:
        try:
            model = next((ln.split(":", 1)[1].strip() for ln in open("/proc/cpuinfo") if ln.startswith("model name")),
                         None)
        except OSError:
            pass
    return {"nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)), "cpu_model": model}


def cpu_threads_default() -> int:
    """All CPUs this process may run on, at most 16: the host CPU share that goes with one GPU."""
    return max(1, min(16, len(os.sched_getaffinity(0))))


def cpu_baseline(scene, threads: int) -> tuple[dict, dict]:
    """The FP64 oracle (the reference's cost structure) runs one full step of the scene from its initial
    state, PCG to the same 1e-6 relative residual, timed on `threads` host threads: one step, not a
    median (a 256^3 step is about 25 s of CPU work on 16 threads, the bound of the sample).  The GPU runs
    the same step on a fresh context, and the two are compared: labels, pressure, post-projection
    divergence, particle positions and velocities."""
    import oracle

    pos = scene.particles()
    prm = scenes.params(transfer=scene.transfer)
    sim = capi.Simulation(scene.grid, prm)
    sim.set_particles(pos)
    st_g = sim.step(scenes.DT)
    g = {"label": sim.get_grid(capi.FIELD_LABEL), "p": sim.get_grid(capi.FIELD_PRESSURE),
         "div": divergence_l2(sim, scene.grid)}
    g["pos"], g["vel"], _ = sim.get_particles()
    sim.close()

    o = oracle.Oracle(scene.grid, prm, threads=threads)
    o.set_particles(pos.astype(np.float64))
    del pos
    t0 = time.perf_counter()
    st_o = o.step(scenes.DT)
    wall = time.perf_counter() - t0
    r = {"label": o.get_grid(capi.FIELD_LABEL), "p": o.get_grid(capi.FIELD_PRESSURE), "div": divergence_l2(o, scene.grid)}
    r["pos"], r["vel"], _ = o.get_particles()
    o.close()

    pn = float(np.linalg.norm(r["p"]))
    acc = {"scope": f"one step of the {scene.n}^3 scene from its initial state, GPU vs the FP64 oracle, "
                    "both with PCG to 1e-6 relative residual",
           "labels_equal": bool(np.array_equal(g["label"], r["label"])),
           "pressure_rel_l2": float(np.linalg.norm(g["p"] - r["p"]) / pn) if pn > 0 else None,
           "divergence_l2_over_b": {"gpu": g["div"] / st_g.b_norm_l2 if st_g.b_norm_l2 > 0 else None,
                                    "oracle": r["div"] / st_o.b_norm_l2 if st_o.b_norm_l2 > 0 else None},
           "residual_rel_l2": {"gpu": st_g.residual_rel_l2, "oracle": st_o.residual_rel_l2},
           "pcg_iterations": {"gpu": st_g.pcg_iterations, "oracle": st_o.pcg_iterations},
           "position_drift_max": float(np.abs(g["pos"] - r["pos"]).max()),
           "position_drift_max_cells": float(np.abs(g["pos"] - r["pos"]).max() / scene.grid.cell_d),
           "velocity_rel_l2": float(np.linalg.norm(g["vel"] - r["vel"]) / max(np.linalg.norm(r["vel"]), 1e-300))}
    host = host_info()
    base = {"value": 1.0 / wall, "unit": "steps/s", "cores": threads, "kind": "port",
            "sample": (f"FP64 oracle (clean-room restatement with the reference's cost structure), ONE full step "
                       f"of the {scene.n}^3 scene from its initial state (not a median): {wall:.1f} s wall with "
                       f"{st_o.pcg_iterations} PCG iterations to 1e-6 on {threads} OpenMP threads"),
            **host}
    return base, acc


def error_trace(steps: int, threads: int) -> dict:
    """The 128^3 APIC dam break with a static sphere run `steps` steps on the GPU and on the FP64 oracle from
    the same initial particles, with no re-synchronisation: per step, each side's relative pressure
    residual ||r||_2/||b||_2 and post-projection divergence L2 / ||b||_2, and the particle-position drift
    between the two (max and RMS, in cells)."""
    import oracle

    sc = scenes.CONFIG_128_APIC_SPHERE()
    prm = scenes.params(transfer=sc.transfer)
    pos = sc.particles()
    sim = capi.Simulation(sc.grid, prm)
    o = oracle.Oracle(sc.grid, prm, threads=threads)
    for side in (sim, o):
        if sc.obstacles:
            side.set_obstacles(sc.obstacles)
    sim.set_particles(pos)
    o.set_particles(pos.astype(np.float64))
    del pos
    rows = []
    t0 = time.perf_counter()
    for k in range(steps):
        st_g, st_o = sim.step(scenes.DT), o.step(scenes.DT)
        pg, _, _ = sim.get_particles()
        po, _, _ = o.get_particles()
        d = np.abs(pg.astype(np.float64) - po) / sc.grid.cell_d
        rows.append({"step": k + 1,
                     "residual_rel_l2": [st_g.residual_rel_l2, st_o.residual_rel_l2],
                     "divergence_l2_over_b": [divergence_l2(sim, sc.grid) / max(st_g.b_norm_l2, 1e-300),
                                              divergence_l2(o, sc.grid) / max(st_o.b_norm_l2, 1e-300)],
                     "drift_max_cells": float(d.max()), "drift_rms_cells": float(np.sqrt(np.mean(d ** 2)))})
        print(f"# error trace step {k + 1}: {rows[-1]}", file=sys.stderr, flush=True)
    sim.close()
    o.close()
    return {"workload": "dam break 128^3 APIC with a static sphere, PCG to 1e-6, GPU and FP64 oracle from the "
                        "same initial particles, never re-synchronised",
            "steps": steps, "oracle_threads": threads, "wall_s": round(time.perf_counter() - t0, 1),
            "columns": "residual_rel_l2 and divergence_l2_over_b are [gpu, oracle]; drift is GPU vs oracle",
            "trace": rows}


COPY_BYTES = 1 << 30   # bytes read (and written) by each calibration copy
PMC_TRAFFIC_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                                "pcg_kernels_256_pmc_traffic.md")


def copy_calibration() -> None:
    """Three plain 1 GiB device-to-device copies: rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
    of this copy kernel give the counter's reading for a known number of bytes."""
    import torch
    a = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": COPY_BYTES, "copies": 3}), flush=True)


def pmc_traffic(kernel: str, size: int) -> float | None:
    """HBM bytes per launch of `kernel` at `size`^3 from the committed rocprofv3 PMC summary (its table row:
    FETCH_SIZE + WRITE_SIZE scaled by the copy-kernel calibration), if the summary covers this run."""
    try:
        lines = open(PMC_TRAFFIC_FILE).read().splitlines()
    except OSError:
        return None
    if not any(ln.strip() == f"- size: {size}" for ln in lines):
        return None
    for ln in lines:
        cells = [c.strip() for c in ln.strip().strip("|").split("|")]
        if len(cells) >= 5 and cells[0].split(" ")[0] == kernel:
            try:
                return float(cells[4].replace(",", ""))
            except ValueError:
                return None
    return None


MFMA_BF16_PEAK_TFLOPS = 2500.0   # MI355X dense BF16


def _state_after(sc, steps: int):
    """Particle state of scene `sc` after `steps` MIC(0)-PCG steps."""
    sim = capi.Simulation(sc.grid, scenes.params(transfer=sc.transfer))
    if sc.obstacles:
        sim.set_obstacles(sc.obstacles)
    sim.set_particles(sc.particles())
    for _ in range(steps):
        sim.step(scenes.DT)
    state = sim.get_particles()
    sim.close()
    return state


def _fpcg_iterations(sc, state, params, budget: int, coarse: int = capi.MG_COARSE_NETWORK):
    """FPCG iterations of one step from `state` on a fresh context (None: not converged within `budget`);
    coarse: how the multigrid's coarsest level is solved (FSIM_OPT_MG_COARSE)."""
    sim = capi.Simulation(sc.grid, scenes.params(transfer=sc.transfer, precond=capi.PRECOND_LEARNED,
                                                 max_iterations=budget))
    if sc.obstacles:
        sim.set_obstacles(sc.obstacles)
    if params is not None:
        sim.learned_set_params(params)
    sim.set_option(capi.OPT_MG_COARSE, coarse)
    sim.set_option(capi.OPT_LEARNED_MIC_COUNT, max(budget // 2, 1))   # FPCG budget 2 x this count
    sim.set_particles(*state)
    st = sim.step(scenes.DT)
    sim.close()
    return st.learned_iterations if st.learned_stop == 0 else None


def _compare_scene(n: int):
    """The n^3 scene of the solver comparison: the 128^3 APIC dam break with a static sphere, the 256^3
    FLIP double dam break of the step-rate measurement, another n: the APIC dam break with a sphere."""
    if n == 128:
        return scenes.CONFIG_128_APIC_SPHERE(), "dam break 128^3 APIC with a static sphere"
    if n == 256:
        return scenes.CONFIG_256_FLIP_DOUBLE(), "double dam break 256^3 FLIP"
    return scenes.dam_break(n, capi.APIC, seed=0, sphere=True), f"dam break {n}^3 APIC with a static sphere"


def learned_solve_compare(sizes, train_n: int, train_scenes: int, epochs: int, seeds: int = 2) -> dict:
    """Pressure-solve ms to 1e-6 relative residual, MIC(0)-PCG vs the learned multigrid FPCG, on the same
    right-hand side.

    The learned preconditioner is a multigrid V-cycle whose coarsest level (16^3 here) is solved by the
    network.  The network is trained here, on the GPU: learned.harvest_samples steps seeded random dam
    breaks at train_n^3 and records the coarsest-level problems of accurate multigrid solves;
    learned.train fits the network to them from `seeds` initial seeds.  Each candidate is validated by its
    FPCG iteration count on a 64^3 FLIP double dam break (another scene than the training's) and the best
    is set on a context of each size n in `sizes`.  Each n^3 scene (_compare_scene) is stepped 3 times
    with MIC(0)-PCG, its particle state is saved, and the next step is solved from that same state with
    each preconditioner (the second of two runs is timed, so nothing is allocated in it).  `learned` is
    the production path: FPCG with the MIC(0)-PCG fallback after twice the last MIC(0) iteration count;
    `mg_smooth` is the same V-cycle with the coarsest level smoothed instead (no network).
    `fpcg_iterations` compares FPCG iteration counts on fresh contexts (budget 10x the MIC(0) count) for
    the coarsest level solved by the trained network, by an untrained network, by smoothing alone and
    accurately (600 sweeps: the bound a perfect network would reach)."""
    t0 = time.perf_counter()
    B, OP, X, dims = learned.harvest_samples(train_n, train_scenes, 3, per_solve=12, seed=0)
    t_gen = time.perf_counter() - t0
    print(f"# learned: {len(B)} coarse samples {dims} from {train_scenes} scenes at {train_n}^3 in {t_gen:.1f}s",
          file=sys.stderr, flush=True)
    val_sc = scenes.double_dam_break(64, capi.FLIP, seed=2)
    val_state = _state_after(val_sc, 3)
    cands = []
    t0 = time.perf_counter()
    for seed in range(1, seeds + 1):
        t1 = time.perf_counter()
        prm, hist = learned.train(dims, B, OP, X, epochs=epochs, batch=8, lr=1e-3, seed=seed)
        t_tr = time.perf_counter() - t1
        its = _fpcg_iterations(val_sc, val_state, prm, 200)
        cands.append({"seed": seed, "params": prm, "loss_first_epoch": hist[0], "loss_last_epoch": hist[-1],
                      "train_s": round(t_tr, 2), "validation_iterations": its})
        print(f"# learned seed {seed}: loss {hist[0]:.3e} -> {hist[-1]:.3e} in {t_tr:.1f}s, "
              f"validation FPCG {its} it", file=sys.stderr, flush=True)
    t_train = time.perf_counter() - t0
    ok = [c for c in cands if c["validation_iterations"] is not None]
    best = min(ok, key=lambda c: c["validation_iterations"]) if ok else cands[0]
    init = capi.Simulation(capi.Grid(*dims, 1.0 / dims[0], 0.25 / dims[0]), scenes.params())
    init.learned_init(best["seed"])
    untrained = init.learned_get_params()
    init.close()
    training = {"samples": len(B), "coarse_grid": list(dims), "scene_grid": [train_n] * 3, "scenes": train_scenes,
                "epochs": epochs, "loss": "masked MSE against the accurate coarse solution",
                "harvest_s": round(t_gen, 2), "train_s": round(t_train, 2),
                "validation": "FPCG iterations (budget 200) of the 4th step of a 64^3 FLIP double dam break",
                "selected_seed": best["seed"],
                "candidates": [{k: v for k, v in c.items() if k != "params"} for c in cands]}
    return {"training": training, "grids": [_compare_at(n, best["params"], untrained) for n in sizes]}, best["params"]


def _compare_at(n: int, params, untrained) -> dict:
    """MIC(0)-PCG vs learned multigrid FPCG on the n^3 comparison scene (learned_solve_compare)."""
    sc, what = _compare_scene(n)
    state = _state_after(sc, 3)
    sim = capi.Simulation(sc.grid, scenes.params(transfer=sc.transfer))
    if sc.obstacles:
        sim.set_obstacles(sc.obstacles)
    sim.learned_set_params(params)
    res = {}

    def solve(name, precond, coarse=capi.MG_COARSE_NETWORK, profile=False):
        sim.set_params(scenes.params(transfer=sc.transfer, precond=precond))
        sim.set_option(capi.OPT_MG_COARSE, coarse)
        for _ in range(2):
            sim.set_particles(*state)
            st = sim.step(scenes.DT)
        res[name] = {"solve_ms": round(st.solve_ms, 3), "pcg_iterations": st.pcg_iterations,
                     "learned_iterations": st.learned_iterations, "learned_stop": st.learned_stop,
                     "fallback": st.fallback, "residual_rel_l2": st.residual_rel_l2}
        print(f"# {n}^3 {name}: {res[name]}", file=sys.stderr, flush=True)
        if not profile:
            return None
        # a third run with per-kernel HIP events: the V-cycle's and the network's time per application
        sim.set_particles(*state)
        sim.set_option(capi.OPT_PROFILE, 1)
        sim.reset_kernel_times()
        sim.step(scenes.DT)
        sim.set_option(capi.OPT_PROFILE, 0)
        return sim.kernel_times()

    solve("mic0_pcg", capi.PRECOND_MIC0)
    kt = solve("learned", capi.PRECOND_LEARNED, profile=True)
    solve("mg_smooth", capi.PRECOND_LEARNED, capi.MG_COARSE_SMOOTH)
    k_mic = res["mic0_pcg"]["pcg_iterations"]
    levels, cdims = sim.mg_info()
    sim.close()
    budget = 10 * k_mic
    iters = {"network_trained": _fpcg_iterations(sc, state, params, budget),
             "network_untrained": _fpcg_iterations(sc, state, untrained, budget),
             "smooth": _fpcg_iterations(sc, state, None, budget, capi.MG_COARSE_SMOOTH),
             "accurate": _fpcg_iterations(sc, state, None, budget, capi.MG_COARSE_ACCURATE),
             "budget": budget, "mic0_pcg": k_mic}
    print(f"# {n}^3 FPCG iterations (None: not converged within the budget): {iters}", file=sys.stderr, flush=True)
    net_ms, net_calls = kt.get("learned", (0.0, 0))
    mg_ms, mg_calls = kt.get("mg", (0.0, 0))
    network = None
    if net_calls > 0 and net_ms > 0:
        fl = learned.flops_per_apply(cdims)
        tf = fl / (net_ms / net_calls * 1e-3) / 1e12
        network = {"bound": "mfma", "achieved": round(tf, 3), "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                   "frac": round(tf / MFMA_BF16_PEAK_TFLOPS, 5), "grid": list(cdims), "flops_per_apply": fl,
                   "ms_per_apply": round(net_ms / net_calls, 4), "applies": net_calls,
                   "note": "the network on the coarsest level: launch-bound at this size, not MFMA-bound"}
    mic_ms, l_ms = res["mic0_pcg"]["solve_ms"], res["learned"]["solve_ms"]
    return {"grid": [n] * 3, "workload": f"{what}, 4th step, same right-hand side for every solver",
            "multigrid": {"levels": levels, "coarsest": list(cdims),
                          "ms_per_vcycle": round(mg_ms / mg_calls, 4) if mg_calls else None,
                          "note": "ms_per_vcycle covers the whole V-cycle, network included; setup counted once"},
            "mic0_pcg": res["mic0_pcg"], "learned": res["learned"], "mg_smooth": res["mg_smooth"],
            "fpcg_iterations": iters,
            "learned_over_mic0_time": round(l_ms / mic_ms, 3) if mic_ms > 0 else None,
            "network": network}


def boxes_source_run(params, steps: int) -> dict:
    """The 256^3 FLIP scene with 3-8 solid boxes and a particle source (scenes.CONFIG_256_BOXES_SOURCE) run
    `steps` steps with MIC(0)-PCG and, on a second context from the same start, with the learned multigrid
    FPCG (the network of the solver comparison) and its MIC(0)-PCG fallback after twice the last MIC(0)
    count; the first learned solve is MIC(0)-PCG, which sets that count, and so is every 50th after.
    Per solver: mean step and solve ms, mean iterations, fallbacks, final particle count."""
    sc = scenes.CONFIG_256_BOXES_SOURCE()
    pos = sc.particles()
    out = {"workload": f"256^3 FLIP double dam break with {len(sc.obstacles) - 1} solid boxes and a particle source "
                       f"(2000 particles per step), {steps} steps from the start, PCG to 1e-6",
           "steps": steps, "initial_particles": len(pos)}
    for name, precond in (("mic0_pcg", capi.PRECOND_MIC0), ("learned", capi.PRECOND_LEARNED)):
        sim = capi.Simulation(sc.grid, scenes.params(transfer=sc.transfer, precond=precond, spawning_enabled=1))
        sim.set_obstacles(sc.obstacles)
        if precond == capi.PRECOND_LEARNED:
            sim.learned_set_params(params)
        sim.set_particles(pos)
        stats = []
        t0 = time.perf_counter()
        for k in range(steps):
            stats.append(sim.step(scenes.DT))
            if k % 50 == 0:
                print(f"# boxes+source {name} step {k}: {stats[-1].step_ms:.1f} ms", file=sys.stderr, flush=True)
        wall = time.perf_counter() - t0
        fpcg = [s.learned_iterations for s in stats if s.learned_stop != 3]
        out[name] = {"ms_per_step": round(1e3 * wall / steps, 3),
                     "solve_ms_mean": round(float(np.mean([s.solve_ms for s in stats])), 3),
                     "pcg_iterations_mean": round(float(np.mean([s.pcg_iterations for s in stats])), 2),
                     "fpcg_iterations_mean": round(float(np.mean(fpcg)), 2) if fpcg else None,
                     "mic0_solves": sum(1 for s in stats if s.learned_stop == 3 or s.fallback),
                     "fallbacks": sum(s.fallback for s in stats),
                     "max_residual_rel_l2": max(s.residual_rel_l2 for s in stats),
                     "final_particles": stats[-1].num_particles}
        print(f"# boxes+source {name}: {out[name]}", file=sys.stderr, flush=True)
        sim.close()
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", type=int, default=256, help="grid cells per axis (default 256)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=cpu_threads_default())
    ap.add_argument("--error-trace-steps", type=int, default=20,
                    help="steps of the 128^3 GPU-vs-oracle error trace (0: skip)")
    ap.add_argument("--profile-steps", type=int, default=None,
                    help="steps timed per kernel with HIP events after the timed region (default: --steps)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain device-to-device copy of known size (for scaling rocprofv3 "
                         "FETCH_SIZE / WRITE_SIZE counters to bytes) and exit")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the pressure-solve comparison MIC(0)-PCG vs learned FPCG")
    ap.add_argument("--learned-sizes", default="128,256",
                    help="comma-separated grid sizes of the MIC(0)-PCG vs learned FPCG comparison")
    ap.add_argument("--learned-train-size", type=int, default=64)
    ap.add_argument("--learned-scenes", type=int, default=24)
    ap.add_argument("--learned-epochs", type=int, default=60)
    ap.add_argument("--boxes-source-steps", type=int, default=50,
                    help="steps of the 256^3 boxes-and-source learned run (the full scene is 500; 0: skip)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported "
              "(the step has no multi-GPU implementation)", file=sys.stderr)
        sys.exit(2)

    scene = scenes.double_dam_break(args.size, capi.FLIP, seed=1)
    grid = scene.grid
    prm = scenes.params(transfer=scene.transfer)
    t_setup = time.perf_counter()
    pos = scene.particles()
    sim = capi.Simulation(grid, prm)
    sim.set_particles(pos)
    n_particles = len(pos)
    del pos
    print(f"# setup {time.perf_counter() - t_setup:.1f}s: {args.size}^3, {n_particles} particles", file=sys.stderr,
          flush=True)
    for k in range(args.warmup):
        st = sim.step(scenes.DT)
        print(f"# warmup {k}: {st.step_ms:.1f} ms, pcg {st.pcg_iterations}", file=sys.stderr, flush=True)

    # timed region: K whole steps (each fsim_step synchronises its stream)
    stats = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        stats.append(sim.step(scenes.DT))
    elapsed = time.perf_counter() - t0
    steps_per_s = args.steps / elapsed
    its = float(np.mean([s.pcg_iterations for s in stats]))
    solve_ms = float(np.mean([s.solve_ms for s in stats]))
    nf = float(np.mean([s.num_fluid_cells for s in stats]))
    print(f"# timed {args.steps} steps in {elapsed:.2f}s: {1e3 * elapsed / args.steps:.1f} ms/step, "
          f"pcg {its:.1f} it, solve {solve_ms:.1f} ms", file=sys.stderr, flush=True)

    # per-kernel HIP-event timing over the same number of further steps
    psteps = args.profile_steps or args.steps
    sim.set_option(capi.OPT_PROFILE, 1)
    sim.reset_kernel_times()
    pstats = [sim.step(scenes.DT) for _ in range(psteps)]
    sim.set_option(capi.OPT_PROFILE, 0)
    kt = sim.kernel_times()
    nf_p = float(np.mean([s.num_fluid_cells for s in pstats]))
    dominant = max(kt, key=lambda k: kt[k][0])
    total_ms, launches = kt[dominant]
    avg_ms = total_ms / max(launches, 1)
    b = algo_bytes(dominant, nf_p, grid.nc, n_particles)
    roofline = None
    if b is not None and avg_ms > 0:
        achieved = b / (avg_ms * 1e-3) / 1e9
        roofline = {"bound": "hbm", "kernel": dominant, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": pmc_traffic(dominant, args.size),
                    "avg_launch_ms": round(avg_ms, 4), "launches": launches,
                    "bytes_per_launch": b}
    breakdown = {k: round(v[0] / psteps, 3) for k, v in sorted(kt.items(), key=lambda kv: -kv[1][0]) if v[1] > 0}
    print(f"# per-step kernel ms: {json.dumps(breakdown)}", file=sys.stderr, flush=True)
    sim.close()

    solve_cmp, boxes = None, None
    if not args.no_learned:
        sizes = [int(v) for v in args.learned_sizes.split(",") if v.strip()]
        solve_cmp, net = learned_solve_compare(sizes, args.learned_train_size, args.learned_scenes, args.learned_epochs)
        if args.boxes_source_steps > 0:
            boxes = boxes_source_run(net, args.boxes_source_steps)

    cpu, accuracy = None, None
    if not args.no_cpu_baseline:
        cpu, accuracy = cpu_baseline(scene, args.cpu_threads)
        print(f"# accuracy vs the FP64 oracle: {json.dumps(accuracy)}", file=sys.stderr, flush=True)

    trace = None
    if args.error_trace_steps > 0 and not args.no_cpu_baseline:
        trace = error_trace(args.error_trace_steps, args.cpu_threads)

    out = {
        "metric": "steps/s @256^3 FLIP",
        "value": round(steps_per_s, 4),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (FP64 dots, residual and pressure)",
        "data": "synthetic",
        "config": {"workload": f"double dam break {args.size}^3 FLIP 0.95, {n_particles} particles, dt 1/240, "
                               "MIC(0)-PCG to 1e-6 relative residual",
                   "grid": [args.size] * 3, "particles": n_particles, "fluid_cells": round(nf),
                   "pcg_iterations": round(its, 1), "pressure_solve_ms": round(solve_ms, 3)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "accuracy": accuracy,
        "error_trace": trace,
        "pressure_solve": solve_cmp,
        "boxes_and_source": boxes,
        "kernel_ms_per_step": breakdown,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
