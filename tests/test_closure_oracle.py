"""Pins of the CPU closure oracle (oracle/closure_oracle.py): the analytic
gradient against central finite differences, the convolution against a
brute-force loop, the transposed convolution as the adjoint of the
convolution, the correction's conservation, Adam's first step, and bf16
rounding on known values."""
import numpy as np
import pytest

from oracle import closure_oracle as C


def rand_layers(rng, scale=0.3):
    return [(rng.normal(size=(co, ci, 3, 3)) * scale, rng.normal(size=co) * 0.1) for ci, co in C.LAYERS]


def test_param_count():
    assert C.num_params() == 21414
    rng = np.random.default_rng(0)
    L = rand_layers(rng)
    back = C.unpack(C.pack(L))
    for (W, b), (W2, b2) in zip(L, back):
        np.testing.assert_array_equal(W, W2)
        np.testing.assert_array_equal(b, b2)


def test_conv_matches_brute_force():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(5, 4, 3))
    W = rng.normal(size=(2, 3, 3, 3))
    b = rng.normal(size=2)
    got = C.conv(x, W, b)
    N, M = x.shape[:2]
    for i in range(N):
        for j in range(M):
            for co in range(2):
                s = b[co]
                for ky in range(3):
                    for kx in range(3):
                        s += x[(i + ky - 1) % N, (j + kx - 1) % M] @ W[co, :, ky, kx]
                assert got[i, j, co] == pytest.approx(s, rel=1e-13, abs=1e-13)


def test_transposed_conv_is_adjoint():
    rng = np.random.default_rng(2)
    x = rng.normal(size=(6, 6, 5))
    g = rng.normal(size=(6, 6, 3))
    W = rng.normal(size=(3, 5, 3, 3))
    lhs = np.sum(C.conv(x, W) * g)
    rhs = np.sum(x * C.conv_input_grad(g, W))
    assert lhs == pytest.approx(rhs, rel=1e-12)
    dW, db = C.conv_weight_grad(g, x)
    # <conv(x, W), g> is linear in W with gradient dW
    assert np.sum(dW * W) == pytest.approx(lhs, rel=1e-12)
    np.testing.assert_allclose(db, g.sum(axis=(0, 1)))


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    layers = rand_layers(rng)
    x0 = rng.normal(size=(6, 6, 4))
    T = rng.normal(size=(6, 6, 4))
    s, w, dx, dy = [0.7, 1.1, 0.9], [1.0, 2.0, 0.5], 0.3, 0.2
    L, g, _ = C.loss_and_grads(layers, x0, T, s, w, dx, dy)
    flat = C.pack(layers)
    h = 1e-6
    for k in rng.choice(flat.size, 60, replace=False):
        fp, fm = flat.copy(), flat.copy()
        fp[k] += h
        fm[k] -= h
        Lp = C.loss_and_grads(C.unpack(fp), x0, T, s, w, dx, dy)[0]
        Lm = C.loss_and_grads(C.unpack(fm), x0, T, s, w, dx, dy)[0]
        fd = (Lp - Lm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * (abs(fd) + abs(g[k])) + 1e-9, (k, fd, g[k])


def test_correction_is_a_flux_divergence():
    rng = np.random.default_rng(4)
    Y = rng.normal(size=(8, 8, 6))
    r = C.correction(Y, [1.0, 2.0, 3.0], 0.1, 0.2)
    np.testing.assert_allclose(r.sum(axis=(0, 1)), 0.0, atol=1e-11)
    # a uniform flux field changes nothing
    assert np.all(C.correction(np.ones((8, 8, 6)), [1, 1, 1], 0.1, 0.1) == 0.0)


def test_adam_first_step_is_lr_times_sign():
    p = np.array([1.0, -2.0, 3.0])
    g = np.array([0.5, -4.0, 1e-3])
    p1, m, v = C.adam(p, g, np.zeros(3), np.zeros(3), 1, 0.01)
    np.testing.assert_allclose(p1, p - 0.01 * np.sign(g), rtol=1e-6)


def test_bf16_rounding():
    assert C.bf16(1.0) == 1.0
    assert C.bf16(1.0 + 2.0 ** -8) == 1.0              # tie rounds to even
    assert C.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6
    assert C.bf16(-3.14159) == -3.140625
