"""FP64 numpy restatement of the learned multigrid preconditioner's V-cycle (TEST INFRASTRUCTURE ONLY).

Only tests/ may import this module, as the checker of the HIP multigrid
(fluid_simulator_amd/csrc/fsim_mg.hip).  The preconditioner has no
counterpart in the reference.  It restates, operation for operation:

  * level 0: the 7-point operator of the interior WATER cells (the pressure
    matrix of calculateAMatrix, bridsonSolverGrid.cpp:40-77, divided by
    dt / (rho dx^2)): diagonal = non-SOLID neighbours, couplings 1 towards
    WATER neighbours; the right-hand side is r / (dt / (rho dx^2));
  * level l+1: 2x2x2 aggregates, piecewise-constant P, R = P^T, Galerkin
    L_{l+1} = P^T L_l P, stored as (diag, cx, cy, cz) per cell;
  * red-black Gauss-Seidel half-sweeps over the parity of x + y + z (odd
    first), 2 sweeps before and 2 after the coarse correction, which is
    added times 1.6; a first sweep from zero sets the other colour to 0;
  * the coarsest level (the first whose largest side is <= 16) is solved by
    `coarse` red-black sweeps from zero (600 for the accurate mode).

Pins (tests/test_mg_oracle.py): the Galerkin operators equal P^T A P built
as sparse matrices; a half-sweep equals the Gauss-Seidel update of the
explicit matrix; the accurate coarse solve equals a direct solve; and the
V-cycle preconditions CG far better than Jacobi.
"""
from __future__ import annotations

import numpy as np

OVER_CORRECT = 1.6
PRE_SWEEPS = POST_SWEEPS = 2
ACCURATE_SWEEPS = 600
COARSE_MAX_SIDE = 16
MAX_LEVELS = 8
WATER, AIR, SOLID = 0, 1, 2


def levels_for(shape) -> int:
    n = list(shape)
    lev = 0
    while max(n) > COARSE_MAX_SIDE and lev < MAX_LEVELS - 1:
        n = [(v + 1) // 2 for v in n]
        lev += 1
    return lev + 1


def fine_operator(label: np.ndarray) -> np.ndarray:
    """op[4][x][y][z] of the interior WATER cells of a label field (SOLID outside the grid)."""
    lab = np.asarray(label)
    shape = lab.shape
    interior = np.zeros(shape, bool)
    interior[1:-1, 1:-1, 1:-1] = True
    fl = (lab == WATER) & interior
    pad = np.full(tuple(n + 2 for n in shape), SOLID, lab.dtype)
    pad[1:-1, 1:-1, 1:-1] = lab
    flp = np.zeros(tuple(n + 2 for n in shape), bool)
    flp[1:-1, 1:-1, 1:-1] = fl
    op = np.zeros((4,) + shape)
    for ax in range(3):
        for d in (-1, 1):
            op[0] += np.roll(pad, -d, axis=ax)[1:-1, 1:-1, 1:-1] != SOLID
        op[1 + ax] = fl & np.roll(flp, -1, axis=ax)[1:-1, 1:-1, 1:-1]
    op[0] = np.where(fl, op[0], 0.0)
    return op


def coarsen(op: np.ndarray) -> np.ndarray:
    """Galerkin operator of the 2x2x2 aggregates: diagonal = children's diagonals less twice the couplings
    inside the aggregate; +a coupling = the couplings of the upper children to the next aggregate."""
    shape = op.shape[1:]
    cs = tuple((n + 1) // 2 for n in shape)
    pad = np.zeros((4,) + tuple(2 * n for n in cs))
    pad[:, :shape[0], :shape[1], :shape[2]] = op
    ch = pad.reshape(4, cs[0], 2, cs[1], 2, cs[2], 2)   # (X, a, Y, b, Z, c) after the field index
    out = np.zeros((4,) + cs)
    out[0] = ch[0].sum(axis=(1, 3, 5))
    for a, axis in enumerate((1, 3, 5)):
        c = ch[1 + a]
        rest = tuple(k if k < axis else k - 1 for k in (1, 3, 5) if k != axis)
        out[0] -= 2.0 * np.take(c, 0, axis=axis).sum(axis=rest)   # lower child -> its sibling
        out[1 + a] = np.take(c, 1, axis=axis).sum(axis=rest)      # upper child -> the next aggregate
    return out


def hierarchy(label: np.ndarray):
    ops = [fine_operator(label)]
    for _ in range(levels_for(label.shape) - 1):
        ops.append(coarsen(ops[-1]))
    return ops


def apply(op: np.ndarray, x: np.ndarray) -> np.ndarray:
    """L x on the unknowns (diag > 0), zero elsewhere."""
    fl = op[0] > 0
    xz = np.where(fl, x, 0.0)
    ax_ = op[0] * xz
    for a in range(3):
        c = op[1 + a]
        up = [slice(None)] * 3
        lo = [slice(None)] * 3
        up[a], lo[a] = slice(1, None), slice(None, -1)
        ax_[tuple(lo)] -= c[tuple(lo)] * xz[tuple(up)]
        ax_[tuple(up)] -= c[tuple(lo)] * xz[tuple(lo)]
    return np.where(fl, ax_, 0.0)


def parity(shape) -> np.ndarray:
    i = np.indices(shape)
    return (i[0] + i[1] + i[2]) % 2


def half_sweep(op, x, b, par: int, first: bool):
    fl = op[0] > 0
    colour = (parity(op.shape[1:]) == par) & fl
    x = x.copy()
    if first:
        s = b
        x[fl & ~colour] = 0.0
    else:
        # b + sum of the couplings times the neighbours = b - (L x - diag x)
        s = b - (apply(op, x) - op[0] * np.where(fl, x, 0.0))
    x[colour] = s[colour] / op[0][colour]
    return x


def smooth(op, x, b, sweeps: int, first: bool):
    for k in range(sweeps):
        x = half_sweep(op, x, b, 1, first and k == 0)
        x = half_sweep(op, x, b, 0, False)
    return x


def restrict(op_f, x, b, shape_c):
    r = np.where(op_f[0] > 0, b - apply(op_f, x), 0.0)
    pad = np.zeros(tuple(2 * n for n in shape_c))
    s = r.shape
    pad[:s[0], :s[1], :s[2]] = r
    return pad.reshape(shape_c[0], 2, shape_c[1], 2, shape_c[2], 2).sum(axis=(1, 3, 5))


def prolong(op_f, x_f, x_c):
    s = x_f.shape
    up = np.repeat(np.repeat(np.repeat(x_c, 2, 0), 2, 1), 2, 2)[:s[0], :s[1], :s[2]]
    return np.where(op_f[0] > 0, x_f + OVER_CORRECT * up, x_f)


def vcycle(ops, r: np.ndarray, scale: float, coarse_sweeps: int = 2, coarse_solve=None) -> np.ndarray:
    """z = one V-cycle of r (on the level-0 grid) for A = scale * L0.  The coarsest level takes
    `coarse_solve(op, b)` if given, else `coarse_sweeps` red-black sweeps from zero."""
    def coarse(op, b):
        if coarse_solve is not None:
            return coarse_solve(op, b)
        return smooth(op, np.zeros_like(b), b, coarse_sweeps, True)

    bs = [np.where(ops[0][0] > 0, r / scale, 0.0)]
    xs = [None] * len(ops)
    c = len(ops) - 1
    if c == 0:
        return coarse(ops[0], bs[0])
    for l in range(c):
        xs[l] = smooth(ops[l], np.zeros_like(bs[l]), bs[l], PRE_SWEEPS, True)
        bs.append(restrict(ops[l], xs[l], bs[l], ops[l + 1].shape[1:]))
        bs[l + 1] = np.where(ops[l + 1][0] > 0, bs[l + 1], 0.0)
    xs[c] = coarse(ops[c], bs[c])
    for l in range(c - 1, -1, -1):
        xs[l] = prolong(ops[l], xs[l], xs[l + 1])
        xs[l] = smooth(ops[l], xs[l], bs[l], POST_SWEEPS, False)
    return xs[0]
