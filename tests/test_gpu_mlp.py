"""The indicator network's device kernels (bf16 matrix-core forward and backward, Adam) against the CPU oracle on the
same bf16-rounded operands: logits and gradients agree to fp32-accumulation rounding (a few differences of one
bfloat16 step in an intermediate), and device training learns what oracle training learns."""
import numpy as np
import pytest

from hcc_hip import MLP
from oracle import mlp as NN

pytestmark = pytest.mark.gpu


def _data(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, NN.F)).astype(np.float32)
    y = (X[:, 0] - 0.5 * X[:, 5] + 0.3 * X[:, 9] ** 2 > 0.4).astype(np.uint8)
    return X, y


@pytest.mark.parametrize("n", [1, 31, 32, 100, 4096])
def test_forward_matches_the_oracle(ctx, n):
    """Logits of n samples (partial and whole 32-sample tiles) for random parameters."""
    X, _ = _data(n, 1)
    rng = np.random.default_rng(2)
    p = (NN.init_params(3) + rng.normal(scale=0.05, size=NN.P)).astype(np.float32)
    m = MLP(ctx, p)
    z = m.forward(X)
    z_o, _ = NN.forward(p.astype(np.float64), X.astype(np.float64), bf16_ops=True)
    assert np.max(np.abs(z - z_o)) <= 2e-2 * (1 + np.max(np.abs(z_o))), np.max(np.abs(z - z_o))
    assert np.median(np.abs(z - z_o)) <= 1e-4 * (1 + np.max(np.abs(z_o)))


@pytest.mark.parametrize("pos_weight", [1.0, 10.0])
def test_gradients_match_the_oracle(ctx, pos_weight):
    """Loss and the gradient of every parameter block on 1000 samples (a partial last tile)."""
    X, y = _data(1000, 4)
    p = NN.init_params(6)
    m = MLP(ctx, p)
    loss, g = m.gradients(X, y, pos_weight)
    loss_o, g_o = NN.loss_and_grad(p.astype(np.float64), X.astype(np.float64), y, pos_weight, bf16_ops=True)
    assert abs(loss - loss_o) <= 1e-4 * abs(loss_o)
    blocks = {"W1": slice(0, 1024), "b1": slice(1024, 1088), "W2": slice(1088, 5184), "b2": slice(5184, 5248),
              "w3": slice(5248, 5312), "b3": slice(5312, 5313)}
    for name, sl in blocks.items():
        err = np.linalg.norm(g[sl] - g_o[sl]) / max(np.linalg.norm(g_o[sl]), 1e-30)
        assert err < 1e-2, (name, err)
    cos = g @ g_o / (np.linalg.norm(g) * np.linalg.norm(g_o))
    assert cos > 0.9999


def test_device_training_learns_like_the_oracle(ctx):
    """Adam on the device over 8192 samples (batch 512, 6 epochs) reaches the oracle's training loss and accuracy."""
    X, y = _data(8192, 7)
    p0 = NN.init_params(8)
    m = MLP(ctx, p0)
    losses = m.train(X, y, epochs=6, batch=512, lr=3e-3, pos_weight=2.0)
    p_o, losses_o = NN.train(p0, X.astype(np.float64), y, epochs=6, batch=512, lr=3e-3, pos_weight=2.0)
    assert losses[-1] < 0.5 * losses[0]
    assert abs(losses[-1] - losses_o[-1]) < 0.1 * losses_o[-1] + 0.01
    Xt, yt = _data(4096, 9)
    acc = np.mean((m.forward(Xt) > 0) == (yt > 0))
    acc_o = np.mean((NN.forward(p_o, Xt.astype(np.float64))[0] > 0) == (yt > 0))
    assert acc > 0.93 and abs(acc - acc_o) < 0.03
