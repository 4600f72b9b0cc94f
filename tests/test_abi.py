"""C-ABI checks that need no GPU: the library loads, exports every function
include/fsim.h declares, and the ctypes structs match the C layout."""
from __future__ import annotations

import ctypes as C
import os
import re
import subprocess
import tempfile

import numpy as np

from fluid_simulator_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fsim.h")


def declared_functions():
    src = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|const char\*)\s+(fsim_\w+)\s*\(", src, re.M)))


def test_library_exports_every_declared_symbol():
    lib = capi.load()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # every declared function has a ctypes prototype in the binding
    assert sorted(capi.PROTOTYPES) == names


def test_status_strings_and_version():
    lib = capi.load()
    assert lib.fsim_abi_version() == 3
    assert lib.fsim_status_string(0) == b"ok"
    assert lib.fsim_status_string(3) == b"unsupported feature"


def _c_layout():
    """sizeof/offsetof of the public structs as the C compiler lays them out."""
    prog = r"""
#include <stdio.h>
#include <stddef.h>
#include "fsim.h"
#define P(T, f) printf(#T "." #f " %zu\n", offsetof(T, f));
int main(void) {
  printf("fsim_grid_desc %zu\n", sizeof(fsim_grid_desc));
  printf("fsim_params %zu\n", sizeof(fsim_params));
  printf("fsim_obstacle %zu\n", sizeof(fsim_obstacle));
  printf("fsim_step_stats %zu\n", sizeof(fsim_step_stats));
  printf("fsim_gfx_particle %zu\n", sizeof(fsim_gfx_particle));
  P(fsim_grid_desc, cell_d) P(fsim_grid_desc, two_d) P(fsim_grid_desc, cell_d_z)
  P(fsim_params, pressure_k) P(fsim_params, max_iterations) P(fsim_params, fluid_density)
  P(fsim_params, stop_rule) P(fsim_params, seed)
  P(fsim_obstacle, pos) P(fsim_obstacle, r) P(fsim_obstacle, last_spawn_fraction)
  P(fsim_step_stats, step_ms) P(fsim_step_stats, residual_abs_inf) P(fsim_step_stats, num_fluid_cells)
  P(fsim_step_stats, solve_nan) P(fsim_step_stats, learned_budget)
  P(fsim_gfx_particle, v) P(fsim_gfx_particle, density)
  return 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "layout.c")
        exe = os.path.join(d, "layout")
        open(c, "w").write(prog)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    return dict(line.rsplit(" ", 1) for line in out.strip().splitlines())


def test_ctypes_structs_match_c_layout():
    lay = {k: int(v) for k, v in _c_layout().items()}
    assert lay["fsim_grid_desc"] == C.sizeof(capi.GridDesc)
    assert lay["fsim_params"] == C.sizeof(capi.Params)
    assert lay["fsim_obstacle"] == C.sizeof(capi.Obstacle)
    assert lay["fsim_step_stats"] == C.sizeof(capi.StepStats)
    assert lay["fsim_gfx_particle"] == C.sizeof(capi.GfxParticle) == capi.GFX_DTYPE.itemsize
    for key, val in lay.items():
        if "." not in key:
            continue
        struct, field = key.split(".")
        cls = {"fsim_grid_desc": capi.GridDesc, "fsim_params": capi.Params, "fsim_obstacle": capi.Obstacle,
               "fsim_step_stats": capi.StepStats, "fsim_gfx_particle": capi.GfxParticle}[struct]
        assert getattr(cls, field).offset == val, key


def test_create_without_device_fails_cleanly():
    """On a machine without a HIP device fsim_create reports NO_DEVICE (and
    never falls back to a CPU path)."""
    from conftest import has_gpu
    if has_gpu():
        return
    lib = capi.load()
    ctx = C.c_void_p()
    g = capi.GridDesc(8, 8, 8, 0.125, 0.03, 0)
    p = capi.make_params()
    assert lib.fsim_create(C.byref(g), C.byref(p), C.byref(ctx)) == capi.ERR_NO_DEVICE
    assert not ctx.value


def test_argument_validation_without_device():
    lib = capi.load()
    ctx = C.c_void_p()
    p = capi.make_params()
    bad = capi.GridDesc(2, 8, 8, 0.125, 0.03, 0)
    assert lib.fsim_create(C.byref(bad), C.byref(p), C.byref(ctx)) == capi.ERR_INVALID
    # the 2D mode: nz = 3 and a z cell size (the depth / 3); the 3D mode: cellD.z == cellD.x
    for nz, two_d, cdz, status in ((3, 1, 0.0, capi.ERR_INVALID), (4, 1, 0.3, capi.ERR_INVALID),
                                   (8, 0, 0.2, capi.ERR_INVALID), (3, 1, 0.3, capi.ERR_NO_DEVICE),
                                   (8, 0, 0.125, capi.ERR_NO_DEVICE)):
        g = capi.GridDesc(8, 8, nz, 0.125, 0.03, two_d, cdz)
        assert lib.fsim_create(C.byref(g), C.byref(p), C.byref(ctx)) == status, (nz, two_d, cdz)
    # the learned multigrid preconditioner is 3D only
    learned = capi.make_params(precond=capi.PRECOND_LEARNED)
    for g, status in ((capi.GridDesc(8, 8, 3, 0.125, 0.03, 1, 0.3), capi.ERR_INVALID),
                      (capi.GridDesc(8, 8, 8, 0.125, 0.03, 0, 0.0), capi.ERR_NO_DEVICE)):
        assert lib.fsim_create(C.byref(g), C.byref(learned), C.byref(ctx)) == status


def test_scene_generator_structure():
    """Host-only generator: 8 particles per fluid cell, one per 2x2x2
    sub-cell, none in border or obstacle cells."""
    grid = capi.Grid(16, 16, 16, 1 / 16, 0.25 / 16)
    sphere = capi.make_obstacle(capi.OBS_SPHERE, (0.25, 0.25, 0.5), r=0.12)
    pos = capi.scene_dam_break(grid, [[0, 0, 0, 0.4, 0.6, 1.0]], [sphere], seed=0)
    cells = np.floor(pos.astype(np.float64) * 16).astype(int)
    sub = np.floor(pos.astype(np.float64) * 32).astype(int) % 2
    key = (cells[:, 0] * 16 + cells[:, 1]) * 16 + cells[:, 2]
    uniq, cnt = np.unique(key, return_counts=True)
    assert np.all(cnt == 8)
    # expected fluid cells: interior, centre inside the block, not in the sphere
    c = np.stack(np.meshgrid(np.arange(16), np.arange(16), np.arange(16), indexing="ij"), -1).reshape(-1, 3)
    ctr = (c + 0.5) / 16
    inside = (ctr[:, 0] < 0.4) & (ctr[:, 1] < 0.6) & np.all((c >= 1) & (c <= 14), axis=1)
    r_cells = 0.12 * 16
    sc = np.array([0.25, 0.25, 0.5]) * 16
    in_sphere = np.sum((c + 0.5 - sc) ** 2, axis=1) < r_cells ** 2
    expect = np.nonzero(inside & ~in_sphere)[0]
    assert np.array_equal(uniq, expect)
    # one particle per sub-cell
    subkey = key * 8 + sub[:, 0] * 4 + sub[:, 1] * 2 + sub[:, 2]
    assert len(np.unique(subkey)) == len(pos)


def test_python_constants_match_the_header():
    """Every enum constant of include/fsim.h that the ctypes binding names has the header's value, and
    every option, stage, field, operator and status code is bound (an enum value added on one side
    only would pass the wrong integer through the ABI)."""
    src = open(HEADER).read()
    consts = {k: int(v) for k, v in re.findall(r"^\s*FSIM_([A-Z0-9_]+)\s*=\s*(-?\d+)", src, re.M)}
    assert len(consts) > 60
    for name, value in consts.items():
        if hasattr(capi, name):
            assert getattr(capi, name) == value, name
    must = [k for k in consts if k.startswith(("OPT_", "STAGE_", "FIELD_", "OP_", "ERR_", "OBS_", "MG_COARSE_"))
            and not k.startswith(("NUM_",))]
    assert "OPT_WAVE_SEND_DELAY" in must
    assert [k for k in must if not hasattr(capi, k)] == []
