"""Pins of the indicator network's CPU oracle: exact-arithmetic gradients against central finite differences, the
bfloat16 rounding against known bit patterns, and training on a separable toy problem."""
import numpy as np
import pytest

from oracle import mlp as NN


def test_bf16_rounding_known_values():
    """Round to nearest, ties to even, on the upper 16 bits of the float."""
    assert NN.bf16_bits([1.0, -2.0, 0.0])[0] == 0x3F80 and NN.bf16_bits([-2.0])[0] == 0xC000
    one_plus = np.float32(1.0) + np.float32(2.0 ** -8)           # exactly halfway between 1 and 1 + 2^-7: to even (1)
    assert NN.bf16(one_plus) == 1.0
    three_half = np.float32(1.0) + np.float32(3 * 2.0 ** -8)     # halfway between 1 + 2^-7 and 1 + 2^-6: to even
    assert NN.bf16(three_half) == 1.0 + 2.0 ** -6
    assert NN.bf16(np.float32(1.0) + np.float32(2.0 ** -9)) == 1.0


@pytest.mark.parametrize("pos_weight", [1.0, 7.5])
def test_exact_gradient_matches_finite_differences(pos_weight):
    """The backward pass of the float64 network equals central differences of the loss (every parameter block)."""
    rng = np.random.default_rng(3)
    X = rng.normal(size=(40, NN.F))
    y = (rng.uniform(size=40) < 0.3).astype(np.uint8)
    p = NN.init_params(5).astype(np.float64) + rng.normal(scale=0.05, size=NN.P)
    _, g = NN.loss_and_grad(p, X, y, pos_weight, bf16_ops=False)
    idx = np.concatenate([rng.choice(NN.F * NN.H, 25, replace=False),                          # W1
                          NN.F * NN.H + rng.choice(NN.H, 5, replace=False),                   # b1
                          NN.F * NN.H + NN.H + rng.choice(NN.H * NN.H, 40, replace=False),    # W2
                          NN.P - 1 - NN.H - NN.H + rng.choice(NN.H, 5, replace=False),        # b2
                          NN.P - 1 - NN.H + rng.choice(NN.H, 10, replace=False), [NN.P - 1]])  # w3, b3
    eps = 1e-6
    for k in idx:
        e = np.zeros(NN.P)
        e[k] = eps
        fd = (NN.loss_and_grad(p + e, X, y, pos_weight, False)[0] - NN.loss_and_grad(p - e, X, y, pos_weight, False)[0]) / (2 * eps)
        assert abs(fd - g[k]) <= 1e-7 + 1e-5 * abs(fd), (k, fd, g[k])


def test_bf16_network_stays_close_to_exact():
    """Rounding the matrix operands to bfloat16 changes logits and gradients at the bfloat16 level only."""
    rng = np.random.default_rng(4)
    X = rng.normal(size=(64, NN.F))
    y = (rng.uniform(size=64) < 0.5).astype(np.uint8)
    p = NN.init_params(1)
    z_e, _ = NN.forward(p, X, False)
    z_b, _ = NN.forward(p, X, True)
    assert np.max(np.abs(z_b - z_e)) < 0.05 * (1 + np.max(np.abs(z_e)))
    _, g_e = NN.loss_and_grad(p, X, y, 1.0, False)
    _, g_b = NN.loss_and_grad(p, X, y, 1.0, True)
    assert np.linalg.norm(g_b - g_e) < 0.05 * np.linalg.norm(g_e)


def test_training_separates_a_toy_problem():
    """Adam on the bf16 network learns a linearly separable label (feature 0 + feature 3 > 0.2)."""
    rng = np.random.default_rng(0)
    X = rng.normal(size=(2048, NN.F))
    y = (X[:, 0] + X[:, 3] > 0.2).astype(np.uint8)
    p, losses = NN.train(NN.init_params(2), X, y, epochs=8, batch=256, lr=3e-3)
    z, _ = NN.forward(p, X)
    assert losses[-1] < 0.5 * losses[0]
    assert np.mean((z > 0) == (y > 0)) > 0.95
