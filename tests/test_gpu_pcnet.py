"""GPU (BF16 MFMA) learned-preconditioner network against the CPU oracle.

Both sides round every convolution operand to bfloat16; the GPU accumulates
in FP32 and the oracle in FP64, so a stored activation can land one bf16 step
apart where the FP32 sum sits on a rounding boundary.  Tolerances: output
1e-3 relative L2, per-layer gradients 1e-2 relative L2.
"""
import numpy as np
import pytest

import mdgpu
from oracle import pcnet as N

pytestmark = pytest.mark.gpu


def _ctx():
    return mdgpu.Context(9, 9, 0.0, 1.0, 0.0, 1.0)


def _data(H, W, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((H, W, 4)).astype(np.float32)
    t = rng.standard_normal((H, W, 3)).astype(np.float32)
    mask = (rng.random((H, W)) > 0.1).astype(np.float32)
    return x, t, mask


def _pad4(a):
    out = np.zeros(a.shape[:2] + (4,), np.float32)
    out[..., :a.shape[2]] = a
    return out


def _mask4(mask):
    m = np.repeat(mask[..., None], 4, axis=2).astype(np.float32)
    m[..., 3] = 0.0
    return m


def rel(a, b):
    return np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-30)


@pytest.mark.parametrize("H,W", [(16, 24), (40, 32)])
def test_forward_matches_oracle(H, W):
    ctx = _ctx()
    P = N.init_params(7)
    for k in range(len(P)):
        P[k][1] = np.random.default_rng(k).normal(0, 0.05, P[k][1].shape).astype(np.float32)
    assert ctx.pcnet_param_count() == N.flatten(P).size
    ctx.pcnet_set_params(N.flatten(P))
    assert np.array_equal(ctx.pcnet_get_params(), N.flatten(P))
    x, _, _ = _data(H, W, 1)
    y = ctx.pcnet_forward(x)
    yo, _ = N.forward(P, x)
    assert rel(y[..., :3], yo) < 1e-3
    assert np.all(y[..., 3] == 0)


def test_gradients_and_adam_match_oracle():
    H, W = 24, 32
    ctx = _ctx()
    P = N.init_params(11)
    ctx.pcnet_set_params(N.flatten(P))
    x, t, mask = _data(H, W, 2)
    loss = ctx.pcnet_grad(x, _pad4(t), _mask4(mask))
    y, A = N.forward(P, x)
    lo, dy = N.loss_grad(y, t, mask)
    G = N.backward(P, A, dy)
    assert loss == pytest.approx(lo, rel=1e-3)
    _, g = ctx.pcnet_get_params(grads=True)
    Gg = N.unflatten(g)
    for k in range(len(P)):
        assert rel(Gg[k][0], G[k][0]) < 1e-2, ("dW", k, rel(Gg[k][0], G[k][0]))
        assert rel(Gg[k][1], G[k][1]) < 1e-2, ("db", k, rel(Gg[k][1], G[k][1]))
    # accumulation doubles the gradient
    ctx.pcnet_grad(x, _pad4(t), _mask4(mask), accumulate=True)
    _, g2 = ctx.pcnet_get_params(grads=True)
    np.testing.assert_allclose(g2, 2 * g, rtol=1e-6, atol=1e-12)
    # one Adam step with the GPU's own gradient: compare with the oracle update
    ctx.pcnet_grad(x, _pad4(t), _mask4(mask))
    _, g = ctx.pcnet_get_params(grads=True)
    ctx.pcnet_adam(1e-3)
    want = N.flatten(N.adam([[w.copy(), b.copy()] for w, b in P],
                            [[gw.astype(np.float64), gb.astype(np.float64)] for gw, gb in N.unflatten(g)],
                            {}, lr=1e-3))
    np.testing.assert_allclose(ctx.pcnet_get_params(), want, rtol=1e-5, atol=1e-7)


def test_training_reduces_loss():
    H, W = 32, 32
    ctx = _ctx()
    ctx.pcnet_set_params(N.flatten(N.init_params(5)))
    rng = np.random.default_rng(3)
    x = rng.standard_normal((H, W, 4)).astype(np.float32)
    # a smooth target that depends on the input: 3x3 box average of channels 1..3
    k = np.ones((3, 3)) / 9.0
    t = np.zeros((H, W, 3), np.float32)
    for c in range(3):
        xp = np.pad(x[..., c + 1], 1)
        t[..., c] = sum(k[i, j] * xp[i:i + H, j:j + W] for i in range(3) for j in range(3))
    m = _mask4(np.ones((H, W)))
    l0 = ctx.pcnet_grad(x, _pad4(t), m)
    for _ in range(150):
        ctx.pcnet_adam(2e-3)
        l1 = ctx.pcnet_grad(x, _pad4(t), m)
    assert l1 < 0.1 * l0, (l0, l1)
