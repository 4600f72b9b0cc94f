"""Learned preconditioner on the GPU (fluid_simulator_amd/csrc/fsim_learned.hip):
the MFMA network against the FP64 oracle (oracle/learned_net.py) with the
same BF16 rounding points, its gradient, training, and the flexible PCG
solve with its MIC(0)-PCG fallback."""
from __future__ import annotations

import numpy as np
import pytest

from fluid_simulator_amd import capi, learned, scenes
from oracle import learned_net as net

pytestmark = pytest.mark.gpu
DT = 1.0 / 240.0


def say(msg):
    print(msg, flush=True)


def solved_sample(n: int, block, seed: int = 0):
    """(b, label, p) of the first pressure solve of a dam break on an n^3 grid (MIC(0)-PCG)."""
    grid = scenes.grid(n)
    pos = capi.scene_dam_break(grid, [block], seed=seed)
    sim = capi.Simulation(grid, scenes.params())
    sim.set_particles(pos)
    st = sim.step(DT)
    b = sim.get_grid(capi.FIELD_RHS)
    lab = sim.get_grid(capi.FIELD_LABEL)
    p = sim.get_grid(capi.FIELD_PRESSURE)
    sim.close()
    return grid, pos, b, lab, p, st


def cosine(a, b):
    a, b = np.asarray(a, np.float64).ravel(), np.asarray(b, np.float64).ravel()
    return float(a @ b / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-300))


def test_forward_matches_oracle(gpu):
    say("case: learned forward 16^3, seed 3")
    grid, _, b, lab, _, _ = solved_sample(16, [0.0, 0.0, 0.0, 0.5, 0.6, 1.0])
    sim = capi.Simulation(grid, scenes.params())
    sim.learned_init(3)
    params = sim.learned_get_params()
    out = sim.learned_forward(b, lab)
    sim.close()
    x, s, fl = net.make_input(b, lab, grid.shape)
    ref, _ = net.forward(net.unpack(params), x, emulate_bf16=True)
    ref = ref[:16, :16, :16].reshape(-1)
    err = np.linalg.norm(out - ref) / np.linalg.norm(ref)
    say(f"  rel L2 error {err:.2e}")
    assert err < 2e-2


def test_lds_convolutions_match_the_cached_kernel_and_the_oracle(gpu):
    # 64^3: level-0 rows of 64 cells and level-1 rows of 32, so the 8->16, 48->16 and 96->32 layers all take
    # the LDS-staged kernel; it sums in the cached kernel's order, so the outputs are bit-identical
    say("case: learned forward 64^3, seed 5, LDS-staged vs cached convolutions")
    grid, _, b, lab, _, _ = solved_sample(64, [0.0, 0.0, 0.0, 0.4, 0.7, 0.9], seed=2)
    sim = capi.Simulation(grid, scenes.params())
    sim.learned_init(5)
    params = sim.learned_get_params()
    out_lds = sim.learned_forward(b, lab)
    sim.set_option(capi.OPT_LEARNED_CONV, 0)
    out_cached = sim.learned_forward(b, lab)
    sim.close()
    assert np.array_equal(out_lds, out_cached)
    x, s, fl = net.make_input(b, lab, grid.shape)
    ref, _ = net.forward(net.unpack(params), x, emulate_bf16=True)
    err = np.linalg.norm(out_lds - ref.reshape(-1)) / np.linalg.norm(ref)
    say(f"  rel L2 error {err:.2e}")
    assert err < 2e-2


def test_gradient_matches_oracle(gpu):
    say("case: learned gradient 16^3, seed 5")
    grid, _, b, lab, p, _ = solved_sample(16, [0.0, 0.0, 0.0, 0.5, 0.6, 1.0])
    sim = capi.Simulation(grid, scenes.params())
    sim.learned_init(5)
    params = sim.learned_get_params()
    loss, g = sim.learned_grad(b, lab, p)
    sim.close()
    x, s, fl = net.make_input(b, lab, grid.shape)
    tgt = np.zeros(x.shape[:3])
    tgt[:16, :16, :16] = np.where(fl, p.reshape(grid.shape) / s, 0.0)
    flp = np.zeros(x.shape[:3], bool)
    flp[:16, :16, :16] = fl
    rloss, rg, _ = net.loss_and_grad(net.unpack(params), x, tgt, flp, emulate_bf16=True)
    say(f"  loss gpu {loss:.6e} oracle {rloss:.6e}")
    assert loss == pytest.approx(rloss, rel=2e-2)
    gp = net.unpack(g)
    for name, _, _ in net.param_layout()[0]:
        c = cosine(gp[name], rg[name])
        rel = np.linalg.norm(gp[name] - rg[name]) / max(np.linalg.norm(rg[name]), 1e-30)
        say(f"  {name}: cosine {c:.5f} rel {rel:.2e}")
        assert c > 0.99 and rel < 0.1, name


def test_energy_gradient_matches_oracle(gpu):
    say("case: learned energy-norm loss gradient 16^3, seed 5")
    grid, _, b, lab, p, _ = solved_sample(16, [0.0, 0.0, 0.0, 0.5, 0.6, 1.0])
    sim = capi.Simulation(grid, scenes.params())
    sim.set_option(capi.OPT_LEARNED_LOSS, 1)
    sim.learned_init(5)
    params = sim.learned_get_params()
    loss, g = sim.learned_grad(b, lab, p)
    sim.close()
    x, s, fl = net.make_input(b, lab, grid.shape)
    tgt = np.zeros(x.shape[:3])
    tgt[:16, :16, :16] = np.where(fl, p.reshape(grid.shape) / s, 0.0)
    flp = np.zeros(x.shape[:3], bool)
    flp[:16, :16, :16] = fl
    op = net.operator_padded(net.label_operator(lab, grid.shape), grid.shape, x.shape[:3])
    rloss, rg, _ = net.energy_loss_and_grad(net.unpack(params), x, tgt, flp, op, emulate_bf16=True)
    say(f"  loss gpu {loss:.6e} oracle {rloss:.6e}")
    assert loss == pytest.approx(rloss, rel=2e-2)
    gp = net.unpack(g)
    for name, _, _ in net.param_layout()[0]:
        c = cosine(gp[name], rg[name])
        rel = np.linalg.norm(gp[name] - rg[name]) / max(np.linalg.norm(rg[name]), 1e-30)
        say(f"  {name}: cosine {c:.5f} rel {rel:.2e}")
        assert c > 0.99 and rel < 0.1, name


def test_training_lowers_the_loss(gpu):
    say("case: learned training 16^3, 2 samples, 40 Adam steps")
    samples = [solved_sample(16, blk) for blk in ([0.0, 0.0, 0.0, 0.5, 0.6, 1.0], [0.3, 0.0, 0.2, 1.0, 0.5, 0.8])]
    grid = samples[0][0]
    B = np.stack([s[2] for s in samples])
    L = np.stack([s[3] for s in samples])
    P = np.stack([s[4] for s in samples])
    sim = capi.Simulation(grid, scenes.params())
    sim.learned_init(7)
    first = sim.learned_train_batch(B, L, P, 1e-3)
    last = first
    for _ in range(39):
        last = sim.learned_train_batch(B, L, P, 1e-3)
    sim.close()
    say(f"  loss {first:.4e} -> {last:.4e}")
    assert np.isfinite(last) and last < 0.5 * first


def test_trained_network_beats_the_solver_without_it(gpu):
    """The network trained on the coarsest-level problems the GPU solver harvests (64^3 random scenes,
    16^3 coarsest level): flexible PCG with it reaches the tolerance without fallback, agrees with
    MIC(0)-PCG, and needs fewer iterations than the same V-cycle without the network (its coarsest level
    left to 2 red-black sweeps) on a scene it was not trained on."""
    say("case: learned multigrid, network trained on harvested 16^3 coarse samples, 64^3 dam break")
    B, OP, X, dims = learned.harvest_samples(64, 3, 2, per_solve=10, seed=11)
    params, hist = learned.train(dims, B, OP, X, epochs=30, batch=8, lr=1e-3, seed=1)
    sc = scenes.dam_break(64, capi.FLIP, sphere=True)
    pos = sc.particles()
    res = {}
    for name, prm, coarse in (("mic0", None, None), ("trained", params, capi.MG_COARSE_NETWORK),
                              ("no_network", None, capi.MG_COARSE_SMOOTH)):
        sim = capi.Simulation(sc.grid, scenes.params(precond=capi.PRECOND_MIC0 if coarse is None
                                                     else capi.PRECOND_LEARNED))
        sim.set_obstacles(sc.obstacles)
        if coarse is not None:
            sim.set_option(capi.OPT_MG_COARSE, coarse)
            sim.set_option(capi.OPT_LEARNED_MIC_COUNT, 500)
        if prm is not None:
            sim.learned_set_params(prm)
        sim.set_particles(pos)
        st = sim.step(DT)
        res[name] = (st, sim.get_grid(capi.FIELD_PRESSURE))
        sim.close()
    st, p = res["trained"]
    st0, _ = res["no_network"]
    st_mic, p_mic = res["mic0"]
    say(f"  {len(B)} samples, loss {hist[0]:.3e} -> {hist[-1]:.3e}; MIC(0)-PCG {st_mic.pcg_iterations} it; "
        f"learned multigrid {st.learned_iterations} it (stop {st.learned_stop}); without the network "
        f"{st0.learned_iterations} it")
    assert hist[-1] < 0.5 * hist[0]
    assert st.fallback == 0 and st.learned_stop == 0
    assert st.residual_rel_l2 <= 1e-6 * 1.01
    # both stop at a 1e-6 relative residual; the pressures differ by at most ~cond(A) times that
    assert np.abs(p - p_mic).max() <= 1e-3 * np.abs(p_mic).max()
    assert st0.fallback == 0
    assert st.learned_iterations < 0.75 * st0.learned_iterations
    assert 2 * st.learned_iterations < st_mic.pcg_iterations


def test_fallback_is_forced_by_an_overflowing_network(gpu):
    say("case: FPCG fallback with parameters 1e30 (the network overflows), 16^3")
    grid, pos, b, lab, p_mic, st_mic = solved_sample(16, [0.0, 0.0, 0.0, 0.5, 0.6, 1.0])
    sim = capi.Simulation(grid, scenes.params(precond=capi.PRECOND_LEARNED))
    sim.learned_set_params(np.full(capi.learned_num_params(), 1e30, np.float32))
    sim.set_option(capi.OPT_LEARNED_MIC_COUNT, st_mic.pcg_iterations)
    sim.set_particles(pos)
    st = sim.step(DT)
    p = sim.get_grid(capi.FIELD_PRESSURE)
    sim.close()
    assert st.fallback == 1 and st.learned_stop == 2 and st.learned_budget == 2 * st_mic.pcg_iterations
    assert st.pcg_iterations == st_mic.pcg_iterations and st.solve_nan == 0
    assert np.array_equal(p, p_mic)


def test_fallback_budget_comes_from_the_context_s_own_mic0_solves(gpu):
    """A fresh context has no MIC(0) count: its first learned solve runs MIC(0)-PCG (learned_stop 3),
    and the FPCG solves after it get twice that count; every FSIM_OPT_LEARNED_REFRESH learned solves
    a MIC(0)-PCG solve renews the count."""
    say("case: FPCG fallback budget from the context's MIC(0) solves, 16^3, refresh period 2")
    grid, pos, b, lab, p_mic, st_mic = solved_sample(16, [0.0, 0.0, 0.0, 0.5, 0.6, 1.0])
    sim = capi.Simulation(grid, scenes.params(precond=capi.PRECOND_LEARNED))
    sim.learned_set_params(np.zeros(capi.learned_num_params(), np.float32))
    sim.set_option(capi.OPT_LEARNED_REFRESH, 2)
    sim.set_particles(pos)
    stats = [sim.step(DT) for _ in range(5)]
    sim.close()
    s0 = stats[0]
    assert s0.learned_stop == 3 and s0.learned_iterations == 0 and s0.fallback == 0
    assert s0.pcg_iterations == st_mic.pcg_iterations
    assert s0.residual_rel_l2 <= 1e-6 * 1.01
    mic, since, calibrations = s0.pcg_iterations, 0, 0
    for st in stats[1:]:
        assert st.residual_rel_l2 <= 1e-6 * 1.01
        if since == 2:   # two learned solves since the last MIC(0) solve
            assert st.learned_stop == 3 and st.learned_iterations == 0
            mic, since, calibrations = st.pcg_iterations, 0, calibrations + 1
            continue
        assert st.learned_stop in (0, 1) and st.learned_budget == 2 * mic
        assert st.learned_iterations <= st.learned_budget
        assert st.fallback == (st.learned_stop == 1)
        since += 1
        if st.fallback:   # the fallback's MIC(0)-PCG solve renews the count too
            mic, since = st.pcg_iterations, 0
    assert calibrations >= 1 or any(st.fallback for st in stats)


def test_learned_without_parameters_is_an_error(gpu):
    grid = scenes.grid(16)
    sim = capi.Simulation(grid, scenes.params(precond=capi.PRECOND_LEARNED))
    sim.set_particles(capi.scene_dam_break(grid, [[0.0, 0.0, 0.0, 0.5, 0.6, 1.0]], seed=0))
    with pytest.raises(capi.FsimError, match="no parameters"):
        sim.step(DT)
    sim.close()
