"""Pins of the learned-preconditioner network oracle: bf16 rounding against
known encodings, the convolution against a direct 9-tap sum, and backward()
against central finite differences of forward() in exact float64."""
import numpy as np
import pytest

from oracle import pcnet as N


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5, 3.0e38, 1e-40],
                 np.float32)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7 -> even (1.0); 1 + 3*2^-8 -> 1 + 2^-6
    want = np.array([1.0, 1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -6, -2.5], np.float32)
    got = N.bf16(x)
    np.testing.assert_array_equal(got[:5], want)
    assert np.all(got.view(np.uint32) & 0xFFFF == 0)


def test_conv_matches_direct_sum():
    rng = np.random.default_rng(0)
    x = N.bf16(rng.standard_normal((6, 5, 4)))
    W = N.bf16(rng.standard_normal((9, 4, 3)))
    b = rng.standard_normal(3)
    y = N.conv(x, W, b)
    ref = np.zeros((6, 5, 3))
    for i in range(6):
        for j in range(5):
            for t in range(9):
                ii, jj = i + t // 3 - 1, j + t % 3 - 1
                if 0 <= ii < 6 and 0 <= jj < 5:
                    ref[i, j] += x[ii, jj].astype(np.float64) @ W[t].astype(np.float64)
    np.testing.assert_allclose(y, ref + b, rtol=1e-13, atol=1e-13)


def test_backward_matches_finite_differences():
    N.ROUND = False
    try:
        rng = np.random.default_rng(1)
        P = N.init_params(3)
        P = [[w.astype(np.float64), b.astype(np.float64) + 0.01 * rng.standard_normal(b.size)]
             for w, b in P]
        x = rng.standard_normal((16, 16, 4))
        t = rng.standard_normal((16, 16, 3))
        mask = (rng.random((16, 16)) > 0.2).astype(np.float64)

        def L(P):
            y, _ = N.forward(P, x)
            return N.loss_grad(y, t, mask)[0]
        y, A = N.forward(P, x)
        _, dy = N.loss_grad(y, t, mask)
        G = N.backward(P, A, dy)
        for k, (ci, co) in enumerate(N.LAYERS):
            for idx in [(0, 0, 0), (4, ci - 1, co - 1), (8, ci // 2, co // 2)]:
                h = 1e-6
                P[k][0][idx] += h
                lp = L(P)
                P[k][0][idx] -= 2 * h
                lm = L(P)
                P[k][0][idx] += h
                fd = (lp - lm) / (2 * h)
                assert G[k][0][idx] == pytest.approx(fd, rel=1e-5, abs=1e-9), (k, idx)
            h = 1e-6
            P[k][1][0] += h
            lp = L(P)
            P[k][1][0] -= 2 * h
            lm = L(P)
            P[k][1][0] += h
            assert G[k][1][0] == pytest.approx((lp - lm) / (2 * h), rel=1e-5, abs=1e-9)
    finally:
        N.ROUND = True


def test_adam_first_step_is_signed_lr():
    P = [[np.ones((9, 1, 1), np.float32), np.zeros(1, np.float32)]]
    G = [[np.full((9, 1, 1), -3.0), np.array([2.0])]]
    st = {}
    N.adam(P, G, st, lr=0.1)
    np.testing.assert_allclose(P[0][0], 1.1, rtol=1e-6)
    np.testing.assert_allclose(P[0][1], -0.1, rtol=1e-6)
