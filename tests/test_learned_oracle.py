"""Pins of the FP64 network oracle (oracle/learned_net.py), CPU only.

The learned preconditioner has no counterpart in the reference, so its oracle
is pinned by mathematics: the data and weight gradients of a convolution are
its adjoints, an impulse reproduces the (mirrored) kernel, pooling inverts
nearest upsampling, and the full backward pass matches central finite
differences of the loss.  A dropped term, a wrong tap sign or a transposed
weight fails one of these."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import learned_net as net


def rnd(rng, *shape, scale=1.0):
    return rng.standard_normal(shape) * scale


def test_param_count_matches_layout():
    n = sum(c * 27 * ci + c for ci, c, _ in net.LAYERS) + net.OUT_IN + 1
    assert net.num_params() == n == 729585


def test_conv3_adjoints():
    rng = np.random.default_rng(0)
    x, w, g = rnd(rng, 4, 6, 5, 3), rnd(rng, 7, 27, 3), rnd(rng, 4, 6, 5, 7)
    y = net.conv3(x, w, None)
    lhs = float((y * g).sum())
    assert lhs == pytest.approx(float((x * net.conv3_dgrad(g, w)).sum()), rel=1e-12)
    assert lhs == pytest.approx(float((w * net.conv3_wgrad(x, g)).sum()), rel=1e-12)


def test_conv3_impulse_is_mirrored_kernel():
    rng = np.random.default_rng(1)
    x = np.zeros((5, 5, 5, 1))
    x[2, 2, 2, 0] = 1.0
    w = rnd(rng, 4, 27, 1)
    y = net.conv3(x, w, None)
    for t, (dx, dy, dz) in enumerate(net.taps()):
        # y[v] = sum_t x[v + off(t)] w[t]: the impulse at c contributes at v = c - off(t)
        assert np.allclose(y[2 - dx, 2 - dy, 2 - dz, :], w[:, t, 0])
    assert np.count_nonzero(np.abs(y).sum(axis=3)) == 27


def test_pool_inverts_up_and_bwd_pairs():
    rng = np.random.default_rng(2)
    x = rnd(rng, 2, 3, 4, 5)
    assert np.allclose(net.pool2(net.up2(x)), x, rtol=1e-15, atol=0)   # sequential FP64 sum of 8 copies
    g = rnd(rng, 4, 6, 8, 5)
    assert float((net.up2(x) * g).sum()) == pytest.approx(float((x * net.up2_bwd(g)).sum()), rel=1e-12)
    assert float((net.pool2(g) * x).sum()) == pytest.approx(float((g * net.pool2_bwd(x)).sum()), rel=1e-12)


def test_bf16_rounding():
    # bf16 keeps 7 fraction bits: the spacing at 1.0 is 2^-7
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7 + 2 ** -8, 1.0 + 2 ** -7 + 2 ** -9, 1.0 + 2 ** -8 + 2 ** -10,
                  -3.14159, 0.0])
    r = net.bf16(v)
    assert r[0] == 1.0
    assert r[1] == 1.0                    # tie between 1 and 1 + 2^-7: to the even one
    assert r[2] == 1.0 + 2 ** -6          # tie between 1 + 2^-7 (odd) and 1 + 2^-6 (even)
    assert r[3] == 1.0 + 2 ** -7          # below the half-way point
    assert r[4] == 1.0 + 2 ** -7          # above the half-way point
    assert r[5] == -3.140625 and r[6] == 0.0


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    shape = (8, 8, 8)
    label = np.ones(shape, np.uint8)     # AIR
    label[1:7, 1:6, 1:7] = 0             # WATER
    b = rng.standard_normal(shape).astype(np.float32)
    x, s, fl = net.make_input(b.reshape(-1), label.reshape(-1), shape)
    target = np.where(fl, rng.standard_normal(shape), 0.0)
    P = {}
    for name, _, shp in net.param_layout()[0]:
        fan = shp[1] * shp[2] if len(shp) == 3 else net.OUT_IN
        P[name] = rnd(rng, *shp, scale=np.sqrt(2.0 / fan)) + (0.05 if name.startswith("b") else 0.0)
    loss, g, _ = net.loss_and_grad(P, x, target, fl)
    h = 1e-6
    checked = 0
    for name, _, shp in net.param_layout()[0]:
        for _ in range(2):
            idx = tuple(int(rng.integers(0, n)) for n in shp)
            old = P[name][idx]
            P[name][idx] = old + h
            lp = net.loss_and_grad(P, x, target, fl)[0]
            P[name][idx] = old - h
            lm = net.loss_and_grad(P, x, target, fl)[0]
            P[name][idx] = old
            fd = (lp - lm) / (2 * h)
            assert fd == pytest.approx(g[name][idx], rel=1e-4, abs=1e-8), (name, idx)
            checked += 1
    assert checked == 2 * len(net.param_layout()[0])


def small_labels(rng, shape=(7, 6, 5)):
    """SOLID border, random WATER / AIR inside."""
    lab = np.full(shape, 2, np.uint8)
    inner = rng.random(tuple(n - 2 for n in shape)) < 0.7
    lab[1:-1, 1:-1, 1:-1] = np.where(inner, 0, 1)
    return lab


def test_energy_form_equals_edge_sum():
    """e^T A e = sum over WATER-WATER edges (e_i - e_j)^2 + sum_i (AIR neighbours of i) e_i^2, by loops."""
    rng = np.random.default_rng(11)
    lab = small_labels(rng)
    shape = lab.shape
    fl = lab == 0
    cnt = net.neighbour_counts(lab.reshape(-1), shape, shape)
    e = np.where(fl, rng.standard_normal(shape), 0.0)
    lhs = float(np.sum(e * net.apply_a(e, fl, cnt)))
    rhs = 0.0
    for x in range(shape[0]):
        for y in range(shape[1]):
            for z in range(shape[2]):
                if not fl[x, y, z]:
                    continue
                for d in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
                    xx, yy, zz = x + d[0], y + d[1], z + d[2]
                    nl = lab[xx, yy, zz] if 0 <= xx < shape[0] and 0 <= yy < shape[1] and 0 <= zz < shape[2] else 2
                    if nl == 1:                                   # AIR: Dirichlet zero
                        rhs += e[x, y, z] ** 2
                    elif nl == 0 and (xx, yy, zz) > (x, y, z):    # each WATER-WATER edge once
                        rhs += (e[x, y, z] - e[xx, yy, zz]) ** 2
    assert lhs == pytest.approx(rhs, rel=1e-12)
    u, v = rng.standard_normal(shape), rng.standard_normal(shape)
    assert float(np.sum(np.where(fl, u, 0) * net.apply_a(v, fl, cnt))) == pytest.approx(
        float(np.sum(np.where(fl, v, 0) * net.apply_a(u, fl, cnt))), rel=1e-12)


def test_energy_loss_matches_finite_differences():
    rng = np.random.default_rng(12)
    shape = (8, 8, 8)
    lab = np.full(shape, 2, np.uint8)
    lab[1:7, 1:7, 1:7] = 1
    lab[1:7, 1:5, 1:7] = 0
    b = rng.standard_normal(shape).astype(np.float32)
    x, s, fl = net.make_input(b.reshape(-1), lab.reshape(-1), shape)
    op = net.operator_padded(net.label_operator(lab.reshape(-1), shape), shape, x.shape[:3])
    target = np.where(fl, rng.standard_normal(shape), 0.0)
    P = {}
    for name, _, shp in net.param_layout()[0]:
        fan = shp[1] * shp[2] if len(shp) == 3 else net.OUT_IN
        P[name] = rnd(rng, *shp, scale=np.sqrt(2.0 / fan)) + (0.05 if name.startswith("b") else 0.0)
    loss, g, _ = net.energy_loss_and_grad(P, x, target, fl, op)
    assert loss > 0
    h = 1e-6
    for name in ("wo", "bo", "w6", "b5", "w3", "w0"):
        shp = P[name].shape
        idx = tuple(int(rng.integers(0, n)) for n in shp)
        old = P[name][idx]
        P[name][idx] = old + h
        lp = net.energy_loss_and_grad(P, x, target, fl, op)[0]
        P[name][idx] = old - h
        lm = net.energy_loss_and_grad(P, x, target, fl, op)[0]
        P[name][idx] = old
        assert (lp - lm) / (2 * h) == pytest.approx(g[name][idx], rel=1e-4, abs=1e-8), (name, idx)


def test_operator_form_equals_label_form_and_is_symmetric():
    """apply_op with the labels' operator equals apply_a (the form the edge sum pins); for arbitrary
    non-negative couplings (a Galerkin coarse operator) apply_op is symmetric, and the input channels
    carry the operator."""
    rng = np.random.default_rng(13)
    lab = small_labels(rng)
    shape = lab.shape
    fl = lab == 0
    cnt = net.neighbour_counts(lab.reshape(-1), shape, shape)
    op = net.label_operator(lab.reshape(-1), shape)
    assert np.array_equal(op[0], np.where(fl, cnt, 0))
    e = rng.standard_normal(shape)
    assert np.allclose(net.apply_op(e, op), net.apply_a(e, fl, cnt), rtol=0, atol=1e-12)
    gop = rng.uniform(0, 4, (4,) + shape)
    gop[0] = np.where(fl, gop[0] + 20, 0.0)
    for a in range(3):
        gop[1 + a] *= fl & np.roll(fl, -1, axis=a)
        last = [slice(None)] * 3
        last[a] = -1
        gop[1 + a][tuple(last)] = 0.0
    u, v = rng.standard_normal(shape), rng.standard_normal(shape)
    assert float(np.sum(np.where(fl, u, 0) * net.apply_op(v, gop))) == pytest.approx(
        float(np.sum(np.where(fl, v, 0) * net.apply_op(u, gop))), rel=1e-12)
    x, s, f = net.make_input_op(e.reshape(-1), gop, shape)
    assert np.allclose(x[:shape[0], :shape[1], :shape[2], 2], gop[0] / 6.0)
    assert np.allclose(x[:shape[0], :shape[1], :shape[2], 3], np.where(fl, gop[1], 0.0))
