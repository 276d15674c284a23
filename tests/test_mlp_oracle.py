"""Pins the CPU oracle of the learned selector (network, optimiser, selection):
finite differences for the gradient, torch's Adam and bfloat16 rounding as
library references, closed-form cases for the features and the top-k rule."""
import numpy as np
import pytest
import torch

import mlp_ref
import select_ref


def _params(rng):
    p = mlp_ref.he_init(rng)
    W1, b1, W2, b2, W3, b3 = mlp_ref.unpack(p)
    return mlp_ref.pack(W1, rng.normal(0, 0.1, mlp_ref.H), W2, rng.normal(0, 0.1, mlp_ref.H), W3, 0.3)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = _params(rng)
    X = rng.random((40, mlp_ref.IN))
    t = rng.normal(0, 1, 40)
    w = rng.random(40)
    w[:5] = 0.0
    _, g = mlp_ref.loss_grad(p, X, t, w, bf16_round=False)
    idx = np.concatenate([rng.choice(np.arange(a, b), 8, replace=False) for a, b in (
        (mlp_ref.OFF_W1, mlp_ref.OFF_B1), (mlp_ref.OFF_B1, mlp_ref.OFF_W2), (mlp_ref.OFF_W2, mlp_ref.OFF_B2),
        (mlp_ref.OFF_B2, mlp_ref.OFF_W3), (mlp_ref.OFF_W3, mlp_ref.OFF_B3))] + [[mlp_ref.OFF_B3]])
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        lp, _ = mlp_ref.loss_grad(p + e, X, t, w, bf16_round=False)
        lm, _ = mlp_ref.loss_grad(p - e, X, t, w, bf16_round=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), "parameter %d: fd %.9g vs analytic %.9g" % (i, fd, g[i])


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(4)
    x = np.concatenate([rng.normal(0, 1, 5000), rng.normal(0, 1e-3, 1000), [0.0, -0.0, 1.0, 1.00390625,
                                                                               1.005859375, 3.0e38]])
    x = x.astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(mlp_ref.bf16(x), ref)


def test_adam_matches_torch():
    rng = np.random.default_rng(5)
    p0 = rng.normal(0, 1, 300)
    tp = torch.tensor(p0, dtype=torch.float64, requires_grad=True)
    opt = torch.optim.Adam([tp], lr=1e-2, betas=(0.9, 0.999), eps=1e-8)
    p, m, v = p0.copy(), np.zeros(300), np.zeros(300)
    for step in range(1, 6):
        g = rng.normal(0, 1, 300)
        opt.zero_grad()
        tp.grad = torch.tensor(g, dtype=torch.float64)
        opt.step()
        p, m, v = mlp_ref.adam(p, g, m, v, step, 1e-2)
        assert np.allclose(p, tp.detach().numpy(), rtol=0, atol=1e-12)


def test_bf16_forward_close_to_float64():
    rng = np.random.default_rng(6)
    p = _params(rng)
    X = rng.random((64, mlp_ref.IN))
    y32, _ = mlp_ref.forward(p, X, bf16_round=True)
    y64, _ = mlp_ref.forward(p, X, bf16_round=False)
    assert np.max(np.abs(y32 - y64)) < 0.05 * (1 + np.max(np.abs(y64)))
    assert not np.array_equal(y32, y64)


def _tiny_panel():
    """N=70 haplotypes, 130 common sites (3 words, the last partial) + 6 rare sites."""
    rng = np.random.default_rng(7)
    N, Lc, Lr = 70, 130, 6
    L = Lc + Lr
    common = np.ones(L, np.uint8)
    rare_sites = np.sort(rng.choice(L, Lr, replace=False))
    common[rare_sites] = 0
    A = rng.random((Lc, N)) < 0.3
    words = np.packbits(np.pad(A, ((0, 0), (0, 128 - N))).reshape(Lc, 2, 64), axis=2,
                        bitorder="little").view(np.uint64).reshape(Lc, 2)
    rp, ri = [0], []
    for l in range(L):
        if not common[l]:
            ri += sorted(rng.choice(N, 2, replace=False).tolist())
        rp.append(len(ri))
    return dict(N=N, L=L, common=common, major=np.zeros(L, np.uint8), common_words=words,
                rare_ptr=np.array(rp, np.int32), rare_idx=np.array(ri, np.int32)), A, rare_sites


def test_features_closed_forms():
    P, A, rare_sites = _tiny_panel()
    L = P["L"]
    c2t = np.nonzero(P["common"])[0]
    H = np.zeros((2, L), np.uint8)
    H[0, c2t] = A[:, 5]                  # H0 = candidate 5 at the common sites
    H[1, c2t] = 1 - A[:, 9]              # H1 = complement of candidate 9
    carrier = P["rare_idx"][P["rare_ptr"][rare_sites[0]]]
    H[0, rare_sites[0]] = 1              # H0 carries a rare allele of `carrier`
    gl = np.zeros((1, L, 3), np.uint8)
    gl[0, :, 1:] = 30                    # confident RR everywhere ...
    flat = np.zeros((1, L), np.uint8)
    flat[0, c2t[:10]] = 1                # ... except 10 sites without data
    pool = np.zeros((1, 4), np.int32)
    pool[0, :3] = [5 | (2 << 20), 9 | (1 << 26) | (1 << 27), carrier]
    site = np.zeros((1, 4), np.int32)
    site[0, :3] = [c2t[70], c2t[0], c2t[129]]   # listed in words 1, 0 and 2 of 3
    X = select_ref.features(P, H, gl, flat, pool, site, np.array([3]), depth=4)
    f5, f9, fc = X[0, 0], X[0, 1], X[0, 2]
    assert f5[0] == np.float32(0.5) and f9[2] == 1 and f9[3] == 1 and f5[2] == 0
    assert f5[1] == 0 and f9[1] == np.float32(1) * (np.float32(1) / np.float32(3))
    # identical to H0: no mismatches anywhere, one run over all 3 words, 1 word back and 2 forward
    assert f5[4] == 0 and f5[6] == 0 and f5[8] == 1 and np.all(f5[10:20:2] == 0)
    assert f5[20] == np.float32(1 / 64) and f5[22] == np.float32(2 / 64)
    # complement of H1: every common site mismatches, no clean word
    assert f9[5] == 1 and f9[7] == 1 and f9[9] == 0 and f9[21] == 0 and f9[23] == 0
    assert f9[11] == 0 and f9[13] == 1 and f9[15] == 0 and f9[17] == 1   # listed in word 0: nothing behind it
    # the segment covers the whole chunk here: alt rate and confident-RR alt counts are whole-chunk counts
    assert f5[27] == np.float32(A[:, 5].sum()) / np.float32(130)
    assert f5[24] == np.float32(A[10:, 5].sum()) / np.float32(120)
    assert f5[30] == np.float32(1) * (np.float32(1) / np.float32(3)) and f5[31] == np.float32(130 / 2048)
    assert fc[28] == 1 and fc[29] == 0 and np.all(X[0, 3] == 0)


def test_topk_rule():
    pool = np.array([[10, 11, 12, 13, 14, 0], [20, 21, 22, 0, 0, 0]], np.int32)
    K = np.array([5, 3], np.int32)
    states = np.zeros_like(pool)
    for t in range(2):
        states[t, :K[t]] = np.sort(pool[t, :K[t]])
    sc = np.log(np.array([[1, 50, 50, 3, 1e-9, 0], [9, 1, 1, 0, 0, 0]], np.float64) * 5 + 1e-3).astype(np.float32)
    st, Ko, cov, kept = select_ref.topk(sc, pool, K, np.array([2, 1]), states, 0.9)
    # ties keep insertion order: 11 and 12 beat 10; target 1: 20 alone holds 9/11 < 0.9 -> pool kept
    assert list(st[0, :2]) == [11, 12] and Ko[0] == 2 and not kept[0]
    assert kept[1] and Ko[1] == 3 and np.array_equal(st[1], states[1])
    assert abs(cov[0] - 100 / 104) < 1e-6 and abs(cov[1] - 9 / 11) < 1e-6
