"""Pins the extrapolator restatement (oracle/extrapolator.py):
- the strided convolution against a direct loop;
- the mode synthesis against a per-sample sum of damped cosines;
- every analytic gradient against central finite differences (float64);
- the bfloat16 rounding against known bit patterns.
"""
import math

import numpy as np

from oracle import extrapolator as X

TINY = X.Dims(L0=32, c=(4, 6, 5), hidden=12, modes=3)


def test_conv_matches_direct_loop():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2, 3, 16))
    W = rng.standard_normal((4, 3, X.KERNEL))
    cols = X.im2col(x)
    y = np.einsum("btk,ok->bot", cols, W.reshape(4, -1))
    ref = np.zeros((2, 4, 8))
    for b in range(2):
        for o in range(4):
            for t in range(8):
                for ci in range(3):
                    for k in range(X.KERNEL):
                        s = X.STRIDE * t + k - X.PAD
                        if 0 <= s < 16:
                            ref[b, o, t] += W[o, ci, k] * x[b, ci, s]
    assert np.allclose(y, ref, rtol=1e-13, atol=1e-13)
    # col2im is the adjoint of im2col
    d = rng.standard_normal(cols.shape)
    assert abs((cols * d).sum() - (x * X.col2im(d, 3, 16)).sum()) < 1e-10


def test_synthesis_is_sum_of_damped_cosines():
    alpha = np.array([[0.01, 0.1]])
    f = np.array([[0.05, 0.2]])
    aV = np.array([[1 + 2j, -0.5j]])
    aI = np.array([[0.3, 1 - 1j]])
    t = X.synth(alpha, f, aV, aI, 20)
    n = np.arange(20)
    V = sum(np.exp(-alpha[0, m] * n) * (aV[0, m].real * np.cos(2 * math.pi * f[0, m] * n)
                                        - aV[0, m].imag * np.sin(2 * math.pi * f[0, m] * n)) for m in range(2))
    assert np.allclose(t[0, 0], V, rtol=1e-12, atol=1e-12)


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(1)
    p = X.init_params(TINY, seed=2)
    for k in p:  # non-zero biases so their gradients are exercised fully
        if k.startswith("b"):
            p[k] = 0.1 * rng.standard_normal(p[k].shape)
    x = rng.standard_normal((3, 2, TINY.L0))
    target = rng.standard_normal((3, 2, 7))
    _, g = X.loss_and_grads(p, x, target, TINY)
    for name in X.ORDER:
        flat = p[name].reshape(-1)
        for q in rng.choice(flat.size, size=min(6, flat.size), replace=False):
            old = flat[q]
            h = 1e-6 * max(1.0, abs(old))
            flat[q] = old + h
            lp, _ = X.loss_and_grads(p, x, target, TINY)
            flat[q] = old - h
            lm, _ = X.loss_and_grads(p, x, target, TINY)
            flat[q] = old
            num = (lp - lm) / (2 * h)
            ana = g[name].reshape(-1)[q]
            assert abs(num - ana) <= 1e-6 * max(1.0, abs(num)), (name, q, num, ana)


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.14159265])
    r = X.bf16(v)
    assert r[0] == 1.0
    assert r[1] == 1.0            # tie to even (mantissa ...0)
    assert r[2] == 1.0 + 2 ** -7  # rounds up
    assert r[3] == -2.5
    assert abs(r[4] - 3.140625) < 1e-12


def test_adam_closed_forms():
    """Adam's bias correction makes the first step lr * g / (|g| + eps), and a
    constant gradient keeps that step size at every t (m-hat = g, v-hat = g^2
    exactly).  A dropped bias correction, a missing square root or a swapped
    beta fails one of these."""
    rng = np.random.default_rng(3)
    g = rng.standard_normal(1000) * np.logspace(-4, 2, 1000)
    w0 = rng.standard_normal(1000)
    lr, eps = 1e-3, 1e-8
    w, m, v = X.adam_step(w0, g, 0.0, 0.0, 1, lr)
    np.testing.assert_allclose(w0 - w, lr * g / (np.abs(g) + eps), rtol=1e-12, atol=0)
    for t in range(2, 40):
        w_prev = w
        w, m, v = X.adam_step(w, g, m, v, t, lr)
        np.testing.assert_allclose(w_prev - w, lr * g / (np.abs(g) + eps), rtol=1e-9, atol=0)
    # a gradient that flips sign: second step has m-hat = (b1 g1 (1-b1) + (1-b1) g2)/(1-b1^2)
    g1, g2 = 2.0, -1.0
    w1, m1, v1 = X.adam_step(0.0, g1, 0.0, 0.0, 1, lr)
    w2, _, _ = X.adam_step(w1, g2, m1, v1, 2, lr)
    mh = (0.9 * 0.1 * g1 + 0.1 * g2) / (1 - 0.81)
    vh = (0.999 * 0.001 * g1 ** 2 + 0.001 * g2 ** 2) / (1 - 0.999 ** 2)
    assert abs((w1 - w2) - lr * mh / (math.sqrt(vh) + eps)) < 1e-15
