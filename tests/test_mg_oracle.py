"""Pins of the numpy restatement of the learned multigrid preconditioner (oracle/mg.py, CPU only).

The preconditioner has no counterpart in the reference; its V-cycle is pinned
by what linear algebra fixes: the Galerkin operators equal P^T A P built as
explicit sparse matrices, a red-black half-sweep is the Gauss-Seidel update
of the explicit matrix restricted to one colour, the accurate coarse solve
converges to a direct solve, and the V-cycle is a far better preconditioner
for CG than Jacobi on a real pressure system."""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import mg


def rand_labels(rng, shape):
    lab = np.full(shape, mg.SOLID, np.uint8)
    inner = rng.random(tuple(n - 2 for n in shape))
    lab[1:-1, 1:-1, 1:-1] = np.where(inner < 0.75, mg.WATER, np.where(inner < 0.93, mg.AIR, mg.SOLID))
    return lab


def matrix(op):
    """The explicit sparse matrix of an operator on its unknowns (dense index = flat order of the unknowns)."""
    shape = op.shape[1:]
    fl = op[0] > 0
    ids = -np.ones(shape, int)
    ids[fl] = np.arange(fl.sum())
    rows, cols, vals = [np.arange(fl.sum())], [np.arange(fl.sum())], [op[0][fl]]
    for a in range(3):
        d = np.zeros(3, int)
        d[a] = 1
        src = np.argwhere(fl & (op[1 + a] != 0))
        dst = src + d
        i, j = ids[tuple(src.T)], ids[tuple(dst.T)]
        v = -op[1 + a][tuple(src.T)]
        rows += [i, j]
        cols += [j, i]
        vals += [v, v]
    n = fl.sum()
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n)), ids


def test_fine_operator_is_the_pressure_matrix():
    rng = np.random.default_rng(1)
    lab = rand_labels(rng, (7, 8, 6))
    op = mg.fine_operator(lab)
    A, ids = matrix(op)
    fl = op[0] > 0
    for x, y, z in np.argwhere(fl)[:40]:
        i = ids[x, y, z]
        row = A.getrow(i).toarray().ravel()
        ns = sum(lab[x + d[0], y + d[1], z + d[2]] != mg.SOLID
                 for d in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)))
        assert row[i] == ns
        for d in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
            q = (x + d[0], y + d[1], z + d[2])
            if fl[q]:
                assert row[ids[q]] == -1.0


def test_galerkin_operators_equal_PtAP():
    rng = np.random.default_rng(2)
    lab = rand_labels(rng, (13, 10, 11))   # odd sides: partial aggregates at the upper ends
    ops = [mg.fine_operator(lab)]
    for _ in range(2):
        ops.append(mg.coarsen(ops[-1]))
    for l in range(2):
        A, ids = matrix(ops[l])
        Ac, idc = matrix(ops[l + 1])
        fl = ops[l][0] > 0
        cells = np.argwhere(fl)
        par = idc[tuple((cells // 2).T)]
        assert np.all(par >= 0)   # every fine unknown has a coarse unknown
        P = sp.csr_matrix((np.ones(len(cells)), (ids[tuple(cells.T)], par)), shape=(A.shape[0], Ac.shape[0]))
        G = (P.T @ A @ P).toarray()
        assert np.allclose(G, Ac.toarray(), rtol=0, atol=1e-12)


def test_half_sweep_is_gauss_seidel_on_one_colour():
    rng = np.random.default_rng(3)
    lab = rand_labels(rng, (9, 9, 9))
    op = mg.coarsen(mg.fine_operator(lab))
    A, ids = matrix(op)
    fl = op[0] > 0
    b = np.where(fl, rng.standard_normal(fl.shape), 0.0)
    x = np.where(fl, rng.standard_normal(fl.shape), 0.0)
    got = mg.half_sweep(op, x, b, 1, False)
    xv, bv = x[fl], b[fl]
    colour = (mg.parity(fl.shape) == 1)[fl]
    D = A.diagonal()
    exp = xv.copy()
    off = A @ xv - D * xv
    exp[colour] = (bv[colour] - off[colour]) / D[colour]   # one colour has no coupling inside itself
    assert np.allclose(got[fl], exp, rtol=0, atol=1e-12)
    assert np.array_equal(got[~fl], x[~fl])


def test_accurate_coarse_solve_converges_to_a_direct_solve():
    rng = np.random.default_rng(4)
    lab = rand_labels(rng, (12, 12, 12))
    op = mg.coarsen(mg.fine_operator(lab))   # a 6^3 level
    A, ids = matrix(op)
    fl = op[0] > 0
    b = np.where(fl, rng.standard_normal(fl.shape), 0.0)
    x = mg.smooth(op, np.zeros_like(b), b, mg.ACCURATE_SWEEPS, True)
    assert np.allclose(x[fl], spla.spsolve(A.tocsc(), b[fl]), rtol=1e-8, atol=1e-10)


def _pcg_iterations(A, b, M, tol=1e-6, maxit=1000):
    x = np.zeros_like(b)
    r = b.copy()
    z = M(r)
    s = z.copy()
    sig = z @ r
    for it in range(1, maxit + 1):
        q = A @ s
        a = sig / (s @ q)
        x += a * s
        r -= a * q
        if np.linalg.norm(r) <= tol * np.linalg.norm(b):
            return it
        zn = M(r)
        # flexible (Polak-Ribiere) beta z'.(r' - r) / z.r: the V-cycle is not symmetric
        sn = zn @ r
        beta = (sn - zn @ (r + a * q)) / sig
        s = zn + beta * s
        sig = sn
    return None


def test_vcycle_preconditions_cg_far_better_than_jacobi():
    """A 32^3 dam-break-like label field (two levels: 32 -> 16): flexible PCG with the V-cycle (coarse level
    by 2 sweeps, no network) needs a fraction of the Jacobi-PCG iterations."""
    n = 32
    lab = np.full((n, n, n), mg.AIR, np.uint8)
    lab[0], lab[-1], lab[:, 0], lab[:, :, 0], lab[:, :, -1] = (mg.SOLID,) * 5
    lab[1:13, 1:19, 1:-1] = mg.WATER
    lab[13:25, 1:6, 1:-1] = mg.WATER
    ops = mg.hierarchy(lab)
    assert len(ops) == 2
    A, ids = matrix(ops[0])
    fl = ops[0][0] > 0
    rng = np.random.default_rng(5)
    b = rng.standard_normal(int(fl.sum()))

    def vc(r):
        g = np.zeros(fl.shape)
        g[fl] = r
        return mg.vcycle(ops, g, 1.0)[fl]

    it_mg = _pcg_iterations(A, b, vc)
    it_j = _pcg_iterations(A, b, lambda r: r / A.diagonal())
    assert it_mg is not None and it_j is not None and it_mg * 4 < it_j
