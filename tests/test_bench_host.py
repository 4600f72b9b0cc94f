"""Host-side pieces of bench.py and fluid_simulator_amd/learned.py (CPU only):
the network's FLOP count, the lookup of the committed PMC traffic table, the
post-projection divergence measure (checked on the FP64 oracle, whose
projected field is divergence-free to the solver tolerance) and the refusal
of more than one GPU."""
from __future__ import annotations

import os
import subprocess
import sys

import numpy as np
import pytest

import bench
from fluid_simulator_amd import learned, scenes
from oracle import learned_net as net

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_network_flops_by_hand():
    # 8^3 grid: levels 8^3, 4^3, 2^3, 1^3 voxels; 2 FLOPs per multiply-add
    v = {0: 512, 1: 64, 2: 8, 3: 1}
    macs = (27 * (8 * 16 * v[0] + 16 * 32 * v[1] + 32 * 64 * v[2] + 64 * 128 * v[3] + 192 * 64 * v[2]
                  + 96 * 32 * v[1] + 48 * 16 * v[0]) + 16 * v[0])
    assert learned.flops_per_apply((8, 8, 8)) == 2 * macs
    # padding to multiples of 8, and the same layer table as the oracle's restatement
    assert learned.flops_per_apply((5, 7, 8)) == learned.flops_per_apply((8, 8, 8))
    assert learned.LAYERS == net.LAYERS


def test_pmc_traffic_lookup():
    fwd, bwd = bench.pmc_traffic("sweep_fwd", 256), bench.pmc_traffic("sweep_bwd", 256)
    assert fwd is not None and bwd is not None and fwd > 1e8 and bwd > 1e8
    assert bench.pmc_traffic("sweep_fwd", 128) is None     # another workload
    assert bench.pmc_traffic("extrapolate", 256) is None   # not in the table
    assert bench.pmc_traffic("p2g", 256) > 1e9


@pytest.mark.parametrize("tol", [1e-10, 1e-4])
def test_divergence_measure_matches_the_solver_residual(tol):
    import oracle
    sc = scenes.dam_break(16)
    o = oracle.Oracle(sc.grid, scenes.params(transfer=sc.transfer, tol=tol), threads=2)
    o.set_particles(sc.particles().astype(np.float64))
    st = o.step(scenes.DT)
    d = bench.divergence_l2(o, sc.grid)
    o.close()
    assert st.b_norm_l2 > 0
    # the divergence of the projected faces is the true residual of the pressure solve
    assert d / st.b_norm_l2 == pytest.approx(st.residual_rel_l2, rel=1e-3, abs=1e-12)
    assert d / st.b_norm_l2 <= tol * 1.01


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr
