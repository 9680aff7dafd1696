"""CPU oracle of the learned preconditioner network (forward, backward, Adam).
TEST INFRASTRUCTURE ONLY.

The network has no counterpart in the reference; this module fixes its
arithmetic so the GPU kernels can be checked operation by operation:

* 4-level U-Net on an H x W image (H, W multiples of 8), NHWC activations.
  conv k: 3x3, stride 1, zero padding, weights W[tap][ci][co] (tap = 3*(dy+1)
  + (dx+1), dy the row offset), bias b[co].
    e0 = relu(c0(x))            4 -> 32   (level 0)
    e1 = relu(c1(pool(e0)))    32 -> 32   (level 1)
    e2 = relu(c2(pool(e1)))    32 -> 64   (level 2)
    e3 = relu(c3(pool(e2)))    64 -> 64   (level 3)
    d2 = relu(c4(e2 + up(e3))) 64 -> 32
    d1 = relu(c5(e1 + up(d2))) 32 -> 32
    y  =      c6(e0 + up(d1))  32 -> 3
  pool = 2x2 mean, up = nearest-neighbour x2.
* Precision: every operand of a convolution (activation, weight, upstream
  gradient) is rounded to bfloat16 (round to nearest even); products are
  accumulated in float64 here (float32 on the GPU); activations are stored
  as bfloat16 after bias and ReLU, gradients in float32.
* Loss 0.5 * sum(mask * (y - t)^2) / sum(mask) over the 3 output channels.
"""
import numpy as np

LAYERS = [(4, 32), (32, 32), (32, 64), (64, 64), (64, 32), (32, 32), (32, 3)]
RELU = [True, True, True, True, True, True, False]
ROUND = True   # False: exact float64 arithmetic (to pin backward() by finite differences)


def bf16(a):
    """Round float32/float64 values to the nearest bfloat16 (ties to even);
    returns float32 values that are exactly representable in bfloat16."""
    if not ROUND:
        return np.asarray(a, np.float64)
    a = np.asarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    return np.where(np.isnan(a), a, out)


def init_params(seed):
    """He-normal weights from a counter-based generator, zero biases."""
    rng = np.random.default_rng(seed)
    P = []
    for ci, co in LAYERS:
        std = np.sqrt(2.0 / (9 * ci))
        P.append([rng.standard_normal((9, ci, co)).astype(np.float32) * np.float32(std),
                  np.zeros(co, np.float32)])
    return P


def _shift(x, dy, dx):
    """out[i, j] = x[i + dy, j + dx] with zeros outside."""
    H, W = x.shape[:2]
    out = np.zeros_like(x)
    ys, ye = max(0, -dy), min(H, H - dy)
    xs, xe = max(0, -dx), min(W, W - dx)
    out[ys:ye, xs:xe] = x[ys + dy:ye + dy, xs + dx:xe + dx]
    return out


def conv(x, W, b):
    """3x3 conv with bf16 operands, float64 accumulation."""
    xb = bf16(x).astype(np.float64)
    Wb = bf16(W).astype(np.float64)
    H, Wd = x.shape[:2]
    y = np.zeros((H, Wd, W.shape[2]))
    for t in range(9):
        dy, dx = t // 3 - 1, t % 3 - 1
        y += _shift(xb, dy, dx) @ Wb[t]
    return y + b


def conv_dx(g, W):
    """Gradient of conv wrt its input: correlation of g with the flipped
    taps and transposed weights (bf16 operands)."""
    gb = bf16(g).astype(np.float64)
    Wb = bf16(W).astype(np.float64)
    H, Wd = g.shape[:2]
    out = np.zeros((H, Wd, W.shape[1]))
    for t in range(9):
        dy, dx = t // 3 - 1, t % 3 - 1
        out += _shift(gb, -dy, -dx) @ Wb[t].T
    return out


def conv_dw(x, g):
    """dW[t][ci][co] = sum_p x[p + off_t][ci] g[p][co] (bf16 operands)."""
    xb = bf16(x).astype(np.float64)
    gb = bf16(g).astype(np.float64)
    ci, co = x.shape[2], g.shape[2]
    dW = np.zeros((9, ci, co))
    G = gb.reshape(-1, co)
    for t in range(9):
        dy, dx = t // 3 - 1, t % 3 - 1
        dW[t] = _shift(xb, dy, dx).reshape(-1, ci).T @ G
    return dW, g.reshape(-1, co).sum(0)


def pool(x):
    H, W, C = x.shape
    return x.reshape(H // 2, 2, W // 2, 2, C).mean(axis=(1, 3))


def pool_bwd(g):
    return np.repeat(np.repeat(g, 2, axis=0), 2, axis=1) / 4.0


def up(x):
    return np.repeat(np.repeat(x, 2, axis=0), 2, axis=1)


def up_bwd(g):
    H, W, C = g.shape
    return g.reshape(H // 2, 2, W // 2, 2, C).sum(axis=(1, 3))


def forward(P, x):
    """Returns y (float64) and the stored activations (bf16-valued)."""
    A = {}

    def layer(k, inp):
        z = conv(inp, *P[k])
        return bf16(np.maximum(z, 0.0)) if RELU[k] else z

    A["x"] = bf16(x)
    A["e0"] = layer(0, A["x"])
    A["p0"] = bf16(pool(A["e0"]))
    A["e1"] = layer(1, A["p0"])
    A["p1"] = bf16(pool(A["e1"]))
    A["e2"] = layer(2, A["p1"])
    A["p2"] = bf16(pool(A["e2"]))
    A["e3"] = layer(3, A["p2"])
    A["s2"] = bf16(A["e2"] + up(A["e3"]))
    A["d2"] = layer(4, A["s2"])
    A["s1"] = bf16(A["e1"] + up(A["d2"]))
    A["d1"] = layer(5, A["s1"])
    A["s0"] = bf16(A["e0"] + up(A["d1"]))
    y = layer(6, A["s0"])
    return y, A


def loss_grad(y, t, mask):
    """Loss value and dL/dy; mask is per pixel (H, W) or per value (H, W, 3)."""
    m = (mask[..., None] if mask.ndim == 2 else mask).astype(np.float64)
    m = np.broadcast_to(m, y.shape)
    n = max(m.sum(), 1)
    d = (y - t) * m
    return 0.5 * float((d * d).sum()) / n, d / n


def backward(P, A, dy):
    """Gradients [dW, db] per layer."""
    G = [None] * 7
    relu_mask = lambda a: (a > 0).astype(np.float64)

    def back(k, inp, dz):
        dW, db = conv_dw(inp, dz)
        G[k] = [dW, db]
        return conv_dx(dz, P[k][0])

    ds0 = back(6, A["s0"], dy)
    de0 = ds0.copy()
    dz5 = up_bwd(ds0) * relu_mask(A["d1"])
    ds1 = back(5, A["s1"], dz5)
    de1 = ds1.copy()
    dz4 = up_bwd(ds1) * relu_mask(A["d2"])
    ds2 = back(4, A["s2"], dz4)
    de2 = ds2.copy()
    dz3 = up_bwd(ds2) * relu_mask(A["e3"])
    dp2 = back(3, A["p2"], dz3)
    de2 += pool_bwd(dp2)
    dz2 = de2 * relu_mask(A["e2"])
    dp1 = back(2, A["p1"], dz2)
    de1 += pool_bwd(dp1)
    dz1 = de1 * relu_mask(A["e1"])
    dp0 = back(1, A["p0"], dz1)
    de0 += pool_bwd(dp0)
    dz0 = de0 * relu_mask(A["e0"])
    back(0, A["x"], dz0)
    return G


def adam(P, G, state, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step in float32 (state: step count and moment lists)."""
    state["t"] = state.get("t", 0) + 1
    t = state["t"]
    if "m" not in state:
        state["m"] = [[np.zeros_like(w), np.zeros_like(b)] for w, b in P]
        state["v"] = [[np.zeros_like(w), np.zeros_like(b)] for w, b in P]
    c1, c2 = 1.0 - b1 ** t, 1.0 - b2 ** t
    for k in range(len(P)):
        for j in range(2):
            g = G[k][j].astype(np.float32)
            m = state["m"][k][j] = np.float32(b1) * state["m"][k][j] + np.float32(1 - b1) * g
            v = state["v"][k][j] = np.float32(b2) * state["v"][k][j] + np.float32(1 - b2) * g * g
            P[k][j] = (P[k][j] - np.float32(lr) * (m / np.float32(c1)) /
                       (np.sqrt(v / np.float32(c2)) + np.float32(eps))).astype(np.float32)
    return P


def flatten(P):
    """Flat parameter vector in the library's order: per layer W then b."""
    return np.concatenate([np.concatenate([w.ravel(), b.ravel()]) for w, b in P]).astype(np.float32)


def unflatten(flat):
    P, o = [], 0
    for ci, co in LAYERS:
        w = flat[o:o + 9 * ci * co].reshape(9, ci, co)
        o += 9 * ci * co
        b = flat[o:o + co]
        o += co
        P.append([w.copy(), b.copy()])
    return P
