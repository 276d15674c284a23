"""GPU parity of the learned selector's network (BF16 MFMA forward/backward,
FP32 accumulation, Adam) against the float64 oracle with the same BF16
rounding points.  Tolerances: outputs and gradients agree to 2e-3 of their
scale (the kernels accumulate in FP32 in another order, which can move a value
across a BF16 rounding boundary now and then)."""
import numpy as np
import pytest

import mlp_ref

pytestmark = pytest.mark.gpu

REL = 2e-3


def _params(rng):
    p = mlp_ref.he_init(rng)
    W1, b1, W2, b2, W3, b3 = mlp_ref.unpack(p)
    return mlp_ref.pack(W1, rng.normal(0, 0.1, mlp_ref.H), W2, rng.normal(0, 0.1, mlp_ref.H), W3, 0.25)


@pytest.fixture(scope="module")
def net():
    from glimpse_amd import hip
    return hip.Mlp(seed=1)


def test_init_is_he_scaled(net):
    p = net.params().astype(np.float64)
    W1, b1, W2, b2, W3, b3 = mlp_ref.unpack(p)
    assert abs(W1.std() - np.sqrt(2 / 32)) < 0.02 and abs(W2.std() - np.sqrt(2 / 128)) < 0.01
    assert np.all(b1 == 0) and np.all(b2 == 0) and b3 == 0


def test_mfma_operand_layout_exact(net):
    """Exactly representable weights: a transposed or shifted MFMA operand changes y."""
    W1 = np.zeros((32, 128))
    for i in range(32):
        W1[i, (5 * i) % 128] = 1.0 + i / 8.0          # distinct, exact in bf16
        W1[i, (5 * i + 64) % 128] = 0.5
    W2 = np.zeros((128, 128))
    for c in range(128):
        W2[c, (3 * c + 1) % 128] = 1.0 + (c % 16) / 16.0
    W3 = np.arange(128) / 64.0 - 1.0
    p = mlp_ref.pack(W1, np.zeros(128), W2, np.zeros(128), W3, 0.0)
    net.set_params(p.astype(np.float32))
    X = np.zeros((96, 32), np.float32)
    for r in range(96):
        X[r, r % 32] = 1.0 + (r // 32) / 4.0
    y = net.forward(X)
    ref, _ = mlp_ref.forward(p, X, bf16_round=True)
    assert np.array_equal(y.astype(np.float64), ref), np.max(np.abs(y - ref))


@pytest.mark.parametrize("n", [1, 63, 64, 1000, 4097])
def test_forward_matches_oracle(net, n):
    rng = np.random.default_rng(10 + n)
    p = _params(rng)
    net.set_params(p.astype(np.float32))
    X = rng.random((n, 32)).astype(np.float32)
    y = net.forward(X)
    p32 = p.astype(np.float32).astype(np.float64)
    ref, _ = mlp_ref.forward(p32, X, bf16_round=True)
    scale = 1.0 + np.max(np.abs(ref))
    assert np.max(np.abs(y - ref)) <= REL * scale
    exact, _ = mlp_ref.forward(p32, X, bf16_round=False)
    assert np.max(np.abs(y - exact)) <= 0.05 * scale   # BF16 operands, FP32 accumulation


@pytest.mark.parametrize("n", [50, 1000, 20000])
def test_backward_matches_oracle(net, n):
    rng = np.random.default_rng(20 + n)
    p = _params(rng)
    net.set_params(p.astype(np.float32))
    X = rng.random((n, 32)).astype(np.float32)
    t = rng.normal(0, 1, n).astype(np.float32)
    w = rng.random(n).astype(np.float32)
    w[rng.random(n) < 0.2] = 0.0
    g, loss = net.backward(X, t, w)
    p32 = p.astype(np.float32).astype(np.float64)
    lref, gref = mlp_ref.loss_grad(p32, X, t, w, bf16_round=True)
    assert abs(loss - lref) <= REL * max(1.0, lref)
    for a, b in ((mlp_ref.OFF_W1, mlp_ref.OFF_B1), (mlp_ref.OFF_B1, mlp_ref.OFF_W2), (mlp_ref.OFF_W2, mlp_ref.OFF_B2),
                 (mlp_ref.OFF_B2, mlp_ref.OFF_W3), (mlp_ref.OFF_W3, mlp_ref.OFF_B3), (mlp_ref.OFF_B3, mlp_ref.NPARAM)):
        d = np.linalg.norm(g[a:b] - gref[a:b]) / max(np.linalg.norm(gref[a:b]), 1e-12)
        assert d <= REL, "gradient block [%d:%d) relative error %.3g" % (a, b, d)
    # deterministic: same batch, same bits
    g2, loss2 = net.backward(X, t, w)
    assert np.array_equal(g, g2) and loss == loss2


def test_adam_step_matches_oracle(net):
    rng = np.random.default_rng(30)
    p = _params(rng).astype(np.float32)
    net.set_params(p)
    X = rng.random((777, 32)).astype(np.float32)
    t = rng.normal(0, 1, 777).astype(np.float32)
    pr, m, v = p.astype(np.float64), np.zeros(p.size), np.zeros(p.size)
    for step in (1, 2, 3):
        g, _ = net.backward(X, t)       # gradient at the current device weights
        net.step(X, t, lr=1e-3)
        pr, m, v = mlp_ref.adam(pr, g.astype(np.float64), m, v, step, 1e-3)
        got = net.params()
        assert np.max(np.abs(got - pr)) <= 1e-6 * (1 + np.max(np.abs(pr)))
        pr = got.astype(np.float64)


def test_training_fits_a_smooth_target(net):
    rng = np.random.default_rng(40)
    net.set_params(mlp_ref.he_init(rng).astype(np.float32))
    X = rng.random((8192, 32)).astype(np.float32)
    t = (np.sin(3 * X[:, 0]) + X[:, 1] * X[:, 2] - 0.5 * X[:, 3]).astype(np.float32)
    first = net.step(X, t, lr=3e-3)
    for _ in range(300):
        last = net.step(X, t, lr=3e-3)
    assert last < 0.05 * first, (first, last)
