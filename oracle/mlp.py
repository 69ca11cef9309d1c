"""CPU oracle of the troubled-node indicator network (test infrastructure only: the tests use it as the checker of
the device kernels, nothing in the product imports it).

MLP 16 -> 64 -> 64 -> 1 with GELU (tanh form): forward, the loss mean_s w_s (softplus(z_s) - y_s z_s) with
w_s = pos_weight where y_s = 1, and its gradient, in float64. With bf16=True the operands of the device's matrix
products are rounded to bfloat16 (round to nearest even) at the same points as the device: inputs, weights, the
layer-1 and layer-2 activations, and the layer-2 and layer-1 pre-activation gradients that enter products
(straight-through: a rounding's derivative is taken as one). With bf16=False the arithmetic is exact float64, which
the finite-difference tests pin. The network is build-defined: the reference has no learned component.

Parameter vector layout (fp32 on the device): W1[16][64], b1[64], W2[64][64], b2[64], w3[64], b3.
"""
import numpy as np

F, H = 16, 64
P = F * H + H + H * H + H + H + 1


def bf16(a):
    """Round to the nearest bfloat16 (ties to even) and back to float64: the device's float -> __bf16 conversion."""
    f = np.asarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def bf16_bits(a):
    """The bfloat16 bit patterns (uint16) of float values: the feature rows of the ABI."""
    f = np.asarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    return (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF).astype(np.uint16)


def unpack(p):
    p = np.asarray(p, np.float64)
    o = 0
    W1 = p[o:o + F * H].reshape(F, H); o += F * H
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    w3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def pack(W1, b1, W2, b2, w3, b3):
    return np.concatenate([np.ravel(W1), b1, np.ravel(W2), b2, w3, [b3]])


def init_params(seed=0):
    """Glorot-uniform weights, zero biases (seeded)."""
    rng = np.random.default_rng(seed)
    W1 = rng.uniform(-1, 1, (F, H)) * np.sqrt(6.0 / (F + H))
    W2 = rng.uniform(-1, 1, (H, H)) * np.sqrt(6.0 / (H + H))
    w3 = rng.uniform(-1, 1, H) * np.sqrt(6.0 / (H + 1))
    return pack(W1, np.zeros(H), W2, np.zeros(H), w3, 0.0).astype(np.float32)


K_GELU = 0.7978845608028654


def gelu(x):
    """GELU, tanh form (the device evaluates tanh as 1 - 2 / (exp(2u) + 1), the same function)."""
    return 0.5 * x * (1.0 + np.tanh(K_GELU * (x + 0.044715 * x ** 3)))


def gelu_grad(x):
    t = np.tanh(K_GELU * (x + 0.044715 * x ** 3))
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * K_GELU * (1.0 + 3.0 * 0.044715 * x * x)


def forward(p, X, bf16_ops=True):
    """Logits z[n] and the intermediates of the backward pass."""
    r = bf16 if bf16_ops else (lambda a: np.asarray(a, np.float64))
    W1, b1, W2, b2, w3, b3 = unpack(p)
    Xb, W1b, W2b, w3b = r(X), r(W1), r(W2), r(w3)
    A1 = Xb @ W1b + b1
    H1 = r(gelu(A1))
    A2 = H1 @ W2b + b2
    H2 = r(gelu(A2))
    z = H2 @ w3b + b3
    return z, (Xb, W2b, w3b, A1, H1, A2, H2)


def loss_and_grad(p, X, y, pos_weight=1.0, bf16_ops=True):
    """(loss, gradient[P]) of mean_s w_s (softplus(z_s) - y_s z_s)."""
    r = bf16 if bf16_ops else (lambda a: np.asarray(a, np.float64))
    y = np.asarray(y, np.float64)
    n = len(y)
    z, (Xb, W2b, w3b, A1, H1, A2, H2) = forward(p, X, bf16_ops)
    w = np.where(y > 0, pos_weight, 1.0)
    loss = float(np.mean(w * (np.logaddexp(0.0, z) - y * z)))
    dz = w * (1.0 / (1.0 + np.exp(-z)) - y) / n
    gw3 = H2.T @ dz
    gb3 = dz.sum()
    d2 = dz[:, None] * w3b[None, :] * gelu_grad(A2)
    gb2 = d2.sum(0)
    d2b = r(d2)
    gW2 = H1.T @ d2b
    d1 = (d2b @ W2b.T) * gelu_grad(A1)
    gb1 = d1.sum(0)
    gW1 = Xb.T @ r(d1)
    return loss, pack(gW1, gb1, gW2, gb2, gw3, gb3)


def train(p, X, y, epochs, batch, lr, pos_weight=1.0, bf16_ops=True):
    """Adam (0.9, 0.999, 1e-8, bias-corrected) over mini-batches in the given order; returns (params, epoch losses)."""
    p = np.asarray(p, np.float64).copy()
    m, v = np.zeros(P), np.zeros(P)
    step, losses = 0, []
    for _ in range(epochs):
        ls = []
        for s0 in range(0, len(y), batch):
            l, g = loss_and_grad(p, X[s0:s0 + batch], y[s0:s0 + batch], pos_weight, bf16_ops)
            step += 1
            m = 0.9 * m + 0.1 * g
            v = 0.999 * v + 0.001 * g * g
            p -= lr * (m / (1 - 0.9 ** step)) / (np.sqrt(v / (1 - 0.999 ** step)) + 1e-8)
            ls.append(l)
        losses.append(float(np.mean(ls)))
    return p, losses
