"""CPU restatement of the learned selector's network and optimiser.

TEST INFRASTRUCTURE ONLY: imported by tests/ (and nothing else), never by the
product package.  The network has no counterpart in the reference.

Network: y = W3 . relu(W2^T relu(W1^T x + b1) + b2) + b3 (32 -> 128 -> 128 -> 1),
parameters in one FP32 block [W1 (32x128), b1, W2 (128x128), b2, W3, b3].
Loss: sum_r w_r (y_r - t_r)^2 / sum_r w_r.

All arithmetic is float64.  With bf16=True, values are rounded to bfloat16
(round to nearest even) where the device kernels round them before a matrix
product: the inputs X, the weights W1 and W2, relu(layer 1) and the two
back-propagated pre-activation gradients.  With bf16=False this is the exact
float64 model, which tests pin with central finite differences.

Adam follows Kingma & Ba (2015), Algorithm 1, with the bias corrections folded
into the step: w -= lr * (m / (1 - b1^t)) / (sqrt(v / (1 - b2^t)) + eps).
"""
import numpy as np

IN, H = 32, 128
OFF_W1 = 0
OFF_B1 = OFF_W1 + IN * H
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H
NPARAM = OFF_B3 + 1


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[OFF_W1:OFF_B1].reshape(IN, H), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(H, H),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3], p[OFF_B3])


def pack(W1, b1, W2, b2, W3, b3):
    return np.concatenate([W1.ravel(), b1, W2.ravel(), b2, W3, [b3]])


def _r(x, on):
    return bf16(x) if on else np.asarray(x, np.float64)


def forward(p, X, bf16_round=True):
    """Returns y and the intermediates the backward pass needs."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    Xb = _r(X, bf16_round)
    W1b, W2b = _r(W1, bf16_round), _r(W2, bf16_round)
    a1 = Xb @ W1b + b1
    H1 = _r(np.maximum(a1, 0.0), bf16_round)
    a2 = H1 @ W2b + b2
    H2 = np.maximum(a2, 0.0)
    y = H2 @ W3 + b3
    return y, (Xb, W1b, W2b, H1, H2)


def loss_grad(p, X, t, w=None, bf16_round=True):
    """Loss and its gradient w.r.t. the parameter block (float64)."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    n = X.shape[0]
    w = np.ones(n) if w is None else np.asarray(w, np.float64)
    y, (Xb, W1b, W2b, H1, H2) = forward(p, X, bf16_round)
    inv = 1.0 / w.sum()
    d = y - np.asarray(t, np.float64)
    loss = float((w * d * d).sum() * inv)
    g = 2.0 * w * d * inv                        # dL/dy
    gW3 = H2.T @ g
    gb3 = g.sum()
    d2 = g[:, None] * W3[None, :] * (H2 > 0)     # dL/d(pre-activation 2)
    gb2 = d2.sum(0)
    D2 = _r(d2, bf16_round)
    gW2 = H1.T @ D2
    d1 = (D2 @ W2b.T) * (H1 > 0)                 # dL/d(pre-activation 1)
    gb1 = d1.sum(0)
    D1 = _r(d1, bf16_round)
    gW1 = Xb.T @ D1
    return loss, pack(gW1, gb1, gW2, gb2, gW3, gb3)


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (step counts from 1); returns new (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    p = p - lr * (m / (1 - b1 ** step)) / (np.sqrt(v / (1 - b2 ** step)) + eps)
    return p, m, v


def he_init(rng):
    W1 = rng.normal(0, np.sqrt(2.0 / IN), (IN, H))
    W2 = rng.normal(0, np.sqrt(2.0 / H), (H, H))
    W3 = rng.normal(0, np.sqrt(1.0 / H), H)
    return pack(W1, np.zeros(H), W2, np.zeros(H), W3, 0.0)
