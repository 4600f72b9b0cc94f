"""FP64 numpy restatement of the learned preconditioner network (TEST INFRASTRUCTURE ONLY).

Only tests/ may import this module, as the checker of the HIP network
(fluid_simulator_amd/csrc/fsim_learned.hip, declared in include/fsim.h).
The network has no counterpart in the reference; this module restates the
architecture documented there:

    in  = (b / s, unknown, diag / 6, cx, cy, cz, 0, 0), s = max |b| over the unknowns   level 0, 8 ch
          (the operator: diagonal diag, couplings -cx, -cy, -cz to the +x, +y, +z neighbour; from a
          label field, the 7-point matrix of its WATER cells)
    e0  = relu(conv3(in, 8->16)); e1 = relu(conv3(pool(e0), 16->32))
    e2  = relu(conv3(pool(e1), 32->64)); e3 = relu(conv3(pool(e2), 64->128))
    d2  = relu(conv3(cat(up(e3), e2), 192->64))
    d1  = relu(conv3(cat(up(d2), e1), 96->32))
    d0  = relu(conv3(cat(up(d1), e0), 48->16))
    out = conv1(d0, 16->1)

conv3: 3x3x3, stride 1, zero padding, weights [cout][27][cin] with tap
t = ((dx+1)*3 + dy+1)*3 + dz+1; pool: 2x2x2 mean; up: nearest.  With
emulate_bf16 the weights, the input and every stored activation are rounded
to bfloat16 (round to nearest even) where the GPU rounds them; sums stay in
FP64 here (FP32 on the GPU).

Training losses: masked MSE of out against p/s over the unknowns, or the
energy norm e^T A e / t^T A t of e = out - p/s (A the sample's operator; for
a label field the unscaled 7-point pressure matrix on WATER cells: diagonal =
number of non-SOLID neighbours, -1 towards WATER neighbours, AIR neighbours
Dirichlet zeros).

Parity pins (tests/test_learned_oracle.py): central finite differences of
both FP64 losses against backward() on every layer, conv3 of a unit impulse
equals the mirrored kernel, pool(up(x)) == x, the 1x1 layer is a dot
product; the energy form equals its edge sum
sum_(WATER-WATER edges) (e_i - e_j)^2 + sum_i (AIR neighbours of i) e_i^2
(loops over cells) and A is symmetric; the operator form apply_op equals that
label form for a label field and is symmetric for any coefficients.
"""
from __future__ import annotations

import numpy as np

LAYERS = [(8, 16, 0), (16, 32, 1), (32, 64, 2), (64, 128, 3), (192, 64, 2), (96, 32, 1), (48, 16, 0)]
IN_CH, OUT_IN = 8, 16


def param_layout():
    """[(name, offset, shape)] in the ABI parameter order."""
    out, o = [], 0
    for i, (cin, cout, _) in enumerate(LAYERS):
        out.append((f"w{i}", o, (cout, 27, cin)))
        o += cout * 27 * cin
        out.append((f"b{i}", o, (cout,)))
        o += cout
    out.append(("wo", o, (OUT_IN,)))
    o += OUT_IN
    out.append(("bo", o, (1,)))
    o += 1
    return out, o


def num_params() -> int:
    return param_layout()[1]


def unpack(params: np.ndarray) -> dict:
    lay, n = param_layout()
    assert params.size == n
    return {name: np.asarray(params[o:o + int(np.prod(shape))], np.float64).reshape(shape) for name, o, shape in lay}


def pack(d: dict) -> np.ndarray:
    lay, n = param_layout()
    out = np.zeros(n, np.float64)
    for name, o, shape in lay:
        out[o:o + int(np.prod(shape))] = d[name].reshape(-1)
    return out


def bf16(x: np.ndarray) -> np.ndarray:
    """Round to bfloat16 (nearest even) through float32, returned as float64."""
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def taps():
    return [(t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1) for t in range(27)]


def shift(x: np.ndarray, dx: int, dy: int, dz: int) -> np.ndarray:
    """y[v] = x[v + (dx,dy,dz)], zero outside (x: [D,H,W,C])."""
    D, H, W = x.shape[:3]
    y = np.zeros_like(x)
    xs = slice(max(0, -dx), min(D, D - dx)), slice(max(0, -dy), min(H, H - dy)), slice(max(0, -dz), min(W, W - dz))
    xd = slice(max(0, dx), min(D, D + dx)), slice(max(0, dy), min(H, H + dy)), slice(max(0, dz), min(W, W + dz))
    y[xs] = x[xd]
    return y


def conv3(x: np.ndarray, w: np.ndarray, b: np.ndarray | None) -> np.ndarray:
    """y[v][co] = b[co] + sum_t sum_ci x[v + off(t)][ci] w[co][t][ci]."""
    y = np.zeros(x.shape[:3] + (w.shape[0],))
    for t, (dx, dy, dz) in enumerate(taps()):
        y += shift(x, dx, dy, dz) @ w[:, t, :].T
    if b is not None:
        y += b
    return y


def conv3_wgrad(x: np.ndarray, g: np.ndarray) -> np.ndarray:
    cin, cout = x.shape[3], g.shape[3]
    dw = np.zeros((cout, 27, cin))
    G = g.reshape(-1, cout)
    for t, (dx, dy, dz) in enumerate(taps()):
        dw[:, t, :] = G.T @ shift(x, dx, dy, dz).reshape(-1, cin)
    return dw


def conv3_dgrad(g: np.ndarray, w: np.ndarray) -> np.ndarray:
    dx_ = np.zeros(g.shape[:3] + (w.shape[2],))
    for t, (dx, dy, dz) in enumerate(taps()):
        dx_ += shift(g @ w[:, t, :], -dx, -dy, -dz)
    return dx_


def pool2(x):
    D, H, W, C = x.shape
    return x.reshape(D // 2, 2, H // 2, 2, W // 2, 2, C).mean(axis=(1, 3, 5))


def up2(x):
    return x.repeat(2, axis=0).repeat(2, axis=1).repeat(2, axis=2)


def up2_bwd(g):
    D, H, W, C = g.shape
    return g.reshape(D // 2, 2, H // 2, 2, W // 2, 2, C).sum(axis=(1, 3, 5))


def pool2_bwd(g):
    return up2(g) / 8.0


def label_operator(label: np.ndarray, shape, water: int = 0, solid: int = 2) -> np.ndarray:
    """The 7-point operator of a label field as op[4][x][y][z] = (diag, cx, cy, cz): diag = non-SOLID
    neighbours of a WATER cell (cells outside the grid count as SOLID), couplings 1 between WATER cells
    and their WATER +x / +y / +z neighbour; 0 elsewhere."""
    lab = label.reshape(shape)
    pad = np.full(tuple(n + 2 for n in shape), solid, lab.dtype)
    pad[1:-1, 1:-1, 1:-1] = lab
    w = lab == water
    op = np.zeros((4,) + tuple(shape))
    for ax in range(3):
        for d in (-1, 1):
            op[0] += np.roll(pad, -d, axis=ax)[1:-1, 1:-1, 1:-1] != solid
        op[1 + ax] = w & (np.roll(pad, -1, axis=ax)[1:-1, 1:-1, 1:-1] == water)
    op[0] = np.where(w, op[0], 0.0)
    return op


def make_input_op(b: np.ndarray, op: np.ndarray, shape):
    """Flat grid-order b and an operator op[4][...] (diag, cx, cy, cz) -> network input [D,H,W,8] (padded to
    multiples of 8): (b / s, unknown, diag / 6, cx, cy, cz, 0, 0), s = max |b| over the unknowns (diag > 0)."""
    nx, ny, nz = shape
    op = np.asarray(op, np.float64).reshape((4,) + tuple(shape))
    fl = op[0] > 0
    bb = np.where(fl, b.reshape(shape).astype(np.float64), 0.0)
    s = float(np.max(np.abs(bb))) if fl.any() else 0.0
    pad = [(0, (-n) % 8) for n in shape]
    x = np.zeros(tuple(n + p[1] for n, p in zip(shape, pad)) + (IN_CH,))
    x[:nx, :ny, :nz, 0] = bb / s if s > 0 else 0.0
    x[:nx, :ny, :nz, 1] = fl
    x[:nx, :ny, :nz, 2] = np.where(fl, op[0] / 6.0, 0.0)
    for a in range(3):
        x[:nx, :ny, :nz, 3 + a] = np.where(fl, op[1 + a], 0.0)
    return x, s, fl


def make_input(b: np.ndarray, label: np.ndarray, shape, water: int = 0):   # FSIM_WATER = 0
    """Flat grid-order b and labels -> network input (make_input_op with the labels' operator)."""
    return make_input_op(b, label_operator(label, shape, water), shape)


def operator_padded(op: np.ndarray, shape, padded) -> np.ndarray:
    out = np.zeros((4,) + tuple(padded))
    out[:, :shape[0], :shape[1], :shape[2]] = np.asarray(op).reshape((4,) + tuple(shape))
    return out


def forward(P: dict, x: np.ndarray, emulate_bf16: bool = False):
    r = bf16 if emulate_bf16 else (lambda a: a)
    W = {k: (r(v) if k.startswith("w") else v) for k, v in P.items()}
    c = {"x0": r(x)}
    e = []
    h = c["x0"]
    for i in range(4):
        if i > 0:
            h = r(pool2(e[i - 1]))
            c[f"p{i - 1}"] = h
        pre = conv3(h, W[f"w{i}"], W[f"b{i}"])
        c[f"pre{i}"] = pre
        e.append(r(np.maximum(pre, 0.0)))
    coarse = e[3]
    d = {}
    for i, lev in ((4, 2), (5, 1), (6, 0)):
        cat = np.concatenate([up2(coarse), e[lev]], axis=3)
        c[f"cat{lev}"] = cat
        pre = conv3(cat, W[f"w{i}"], W[f"b{i}"])
        c[f"pre{i}"] = pre
        d[lev] = r(np.maximum(pre, 0.0))
        coarse = d[lev]
    out = d[0] @ W["wo"] + W["bo"][0]
    c.update({"e": e, "d": d, "W": W})
    return out, c


def backward(c: dict, dout: np.ndarray) -> dict:
    """Gradients of sum(dout * out) with respect to every parameter."""
    W, e, d = c["W"], c["e"], c["d"]
    g = {"wo": np.tensordot(dout, d[0], axes=([0, 1, 2], [0, 1, 2])), "bo": np.array([dout.sum()])}
    dy = dout[..., None] * W["wo"]
    dskip = {}
    for i, lev in ((6, 0), (5, 1), (4, 2)):
        gi = dy * (c[f"pre{i}"] > 0)
        g[f"w{i}"] = conv3_wgrad(c[f"cat{lev}"], gi)
        g[f"b{i}"] = gi.sum(axis=(0, 1, 2))
        dcat = conv3_dgrad(gi, W[f"w{i}"])
        cs = LAYERS[i][0] // 3
        cu = LAYERS[i][0] - cs
        dskip[lev] = dcat[..., cu:]
        dy = up2_bwd(dcat[..., :cu])
    # bottleneck and encoder
    de = dy
    for i in (3, 2, 1, 0):
        if i < 3:
            de = dskip[i] + pool2_bwd(dpool)
        gi = de * (c[f"pre{i}"] > 0)
        xin = c["x0"] if i == 0 else c[f"p{i - 1}"]
        g[f"w{i}"] = conv3_wgrad(xin, gi)
        g[f"b{i}"] = gi.sum(axis=(0, 1, 2))
        if i > 0:
            dpool = conv3_dgrad(gi, W[f"w{i}"])
    return g


def loss_and_grad(P: dict, x: np.ndarray, target: np.ndarray, fluid: np.ndarray, emulate_bf16: bool = False):
    """Masked MSE: sum over WATER cells of (out - target)^2 / n_water, and its parameter gradient."""
    out, c = forward(P, x, emulate_bf16)
    n = max(int(fluid.sum()), 1)
    err = np.where(fluid, out - target, 0.0)
    loss = float((err ** 2).sum() / n)
    g = backward(c, 2.0 * err / n)
    return loss, g, out


def neighbour_counts(label: np.ndarray, shape, padded, water: int = 0, solid: int = 2) -> np.ndarray:
    """Non-SOLID neighbour count of each WATER cell (0 elsewhere) on the padded grid; cells outside the
    grid count as SOLID."""
    lab = label.reshape(shape)
    pad = np.full(tuple(n + 2 for n in shape), solid, lab.dtype)
    pad[1:-1, 1:-1, 1:-1] = lab
    cnt = np.zeros(shape, np.int64)
    for ax in range(3):
        for d in (-1, 1):
            cnt += np.roll(pad, -d, axis=ax)[1:-1, 1:-1, 1:-1] != solid
    cnt = np.where(lab == water, cnt, 0)
    out = np.zeros(padded, np.int64)
    out[:shape[0], :shape[1], :shape[2]] = cnt
    return out


def apply_a(e: np.ndarray, fluid: np.ndarray, cnt: np.ndarray) -> np.ndarray:
    """A e on WATER cells (e is taken as zero off WATER), zero elsewhere."""
    ez = np.where(fluid, e, 0.0)
    nb = np.zeros_like(ez)
    for ax in range(3):
        for d in (-1, 1):
            sh = np.zeros_like(ez)
            src = [slice(None)] * 3
            dst = [slice(None)] * 3
            if d > 0:
                src[ax], dst[ax] = slice(1, None), slice(None, -1)
            else:
                src[ax], dst[ax] = slice(None, -1), slice(1, None)
            sh[tuple(dst)] = ez[tuple(src)]
            nb += sh
    return np.where(fluid, cnt * ez - nb, 0.0)


def apply_op(e: np.ndarray, op: np.ndarray) -> np.ndarray:
    """A e for the operator op[4][...] (diagonal op[0], couplings -op[1..3] between a cell and its +x, +y,
    +z neighbour, symmetric) on its unknowns (op[0] > 0; e is taken as zero elsewhere)."""
    fl = op[0] > 0
    ez = np.where(fl, e, 0.0)
    ae = op[0] * ez
    for ax in range(3):
        c = op[1 + ax]
        up = [slice(None)] * 3
        lo = [slice(None)] * 3
        up[ax], lo[ax] = slice(1, None), slice(None, -1)
        ae[tuple(lo)] -= c[tuple(lo)] * ez[tuple(up)]    # the +ax neighbour
        ae[tuple(up)] -= c[tuple(lo)] * ez[tuple(lo)]    # the -ax neighbour (the coupling stored below it)
    return np.where(fl, ae, 0.0)


def energy_loss_and_grad(P: dict, x: np.ndarray, target: np.ndarray, fluid: np.ndarray, op: np.ndarray,
                         emulate_bf16: bool = False):
    """Energy norm e^T A e / t^T A t of e = out - target on the unknowns, A the operator op[4][...] on the
    network's (padded) grid, and its parameter gradient."""
    out, c = forward(P, x, emulate_bf16)
    e = np.where(fluid, out - target, 0.0)
    ae = apply_op(e, op)
    et = float(np.sum(np.where(fluid, target, 0.0) * apply_op(target, op)))
    loss = float(np.sum(e * ae) / et)
    g = backward(c, 2.0 * ae / et)
    return loss, g, out
