"""GPU parity of the learned selector's data path: features (bit-exact with the
numpy restatement), top-k with the mass fallback (exact), training rows
recorded by a run (labels are a probability distribution per target and
iteration), and a learned run that keeps k states where it is confident."""
import numpy as np
import pytest

import select_ref

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sel_case():
    from glimpse_amd import hip, synth
    P = synth.make_panel(300, 1500, 3, length_mb=0.5, rare_frac=0.3, flip_frac=0.05, seed=21)
    gl, flat = synth.make_reads(P["tgt_haps"], 1.0, seed=22)
    return P, gl, flat, hip.Panel(P, with_pbwt=False)


def _pools(rng, T, N, ks, depth):
    K = rng.integers(1, ks + 1, T).astype(np.int32)
    K[0] = ks
    pool = np.zeros((T, ks), np.int32)
    for t in range(T):
        ids = rng.choice(N, K[t], replace=False)
        codes = ids | (rng.integers(0, depth, K[t]) << 20) | (rng.integers(0, 2, K[t]) << 26) | \
            (rng.integers(0, 2, K[t]) << 27)
        pool[t, :K[t]] = codes
    return pool, K


def test_features_bit_exact(sel_case):
    P, gl, flat, GP = sel_case
    rng = np.random.default_rng(1)
    T = gl.shape[0]
    H = P["tgt_haps"].copy()
    H ^= (rng.random(H.shape) < 0.02).astype(np.uint8)   # sampled haplotypes are noisy copies
    pool, K = _pools(rng, T, P["N"], 200, 12)
    common_sites = np.nonzero(P["common"])[0]
    site = rng.choice(common_sites, pool.shape).astype(np.int32)
    site[:, 0] = common_sites[0]
    site[:, 1] = common_sites[-1]
    got = GP.select_features(H, gl, flat, pool, site, K, depth=12)
    ref = select_ref.features(P, H, gl, flat, pool, site, K, 12)
    for t in range(T):
        bad = np.argwhere(got[t] != ref[t])
        assert bad.size == 0, "target %d: %d features differ, first at %s: %r vs %r" % (
            t, len(bad), bad[0], got[t][tuple(bad[0])], ref[t][tuple(bad[0])])


@pytest.mark.parametrize("min_mass", [0.0, 0.95])
def test_topk_exact(min_mass):
    from glimpse_amd import hip
    rng = np.random.default_rng(2)
    T, ks = 12, 700
    pool, K = _pools(rng, T, 100000, ks, 12)
    K[1] = 1
    states = np.zeros_like(pool)
    for t in range(T):
        states[t, :K[t]] = np.sort(pool[t, :K[t]] & 0xFFFFF)
    sc = rng.normal(0, 2, (T, ks)).astype(np.float32)
    sc[2, :50] = 1.5                      # ties
    sc[3] = np.log(1e-3) - 0.01           # no predicted mass at all -> pool kept
    sc[4, :K[4]] = -20.0
    sc[4, 7] = 5.0                        # one dominant candidate
    kl = rng.integers(1, ks, T).astype(np.int32)
    kl[5] = 10 * ks
    st, Ko, cov = hip.select_topk(sc, pool, K, kl, states, min_mass)
    rst, rKo, rcov, kept = select_ref.topk(sc, pool, K, kl, states, min_mass)
    assert np.array_equal(Ko, rKo)
    for t in range(T):
        assert np.array_equal(st[t, :Ko[t]], rst[t, :rKo[t]]), t
    assert np.allclose(cov, rcov, atol=1e-5)


@pytest.fixture(scope="module")
def trained(small_case):
    """A network trained on rows a plain run recorded (the training half)."""
    from glimpse_amd import hip
    P, gl, flat = small_case
    panel = hip.Panel(P)
    prm = hip.default_params(burnin=3, main_iter=4, kpbwt=400)
    run = hip.Run(panel, gl[:2], flat[:2], prm)
    run.collect(1)
    run.execute()
    X, y, w = run.collected()
    net = hip.Mlp(seed=3)
    losses = net.train_run(run, epochs=6, batch=2048, lr=2e-3, seed=4)
    return panel, net, run, (X, y, w), losses


def test_collected_rows_are_distributions(trained):
    panel, net, run, (X, y, w), losses = trained
    st = run.stats()
    n_iter = run.n_iter
    ks = 400
    assert X.shape[0] == st["collected_rows"] == 2 * ks * (n_iter - 1)
    assert np.all(np.isfinite(X)) and np.all(np.isfinite(y))
    groups, wts, feats = y.reshape(-1, ks), w.reshape(-1, ks), X.reshape(-1, ks, 32)
    for g, ww, f in zip(groups, wts, feats):
        K = int(ww.sum())
        assert K > 0 and set(np.unique(ww)) <= {0.0, 1.0} and np.all(ww[K:] == 0) if K < ks else K > 0
        # labels are ln(K m + eps) with m a candidate's average posterior mass over its segment
        m = (np.exp(g[ww > 0].astype(np.float64)) - select_ref.LABEL_EPS) / K
        assert m.min() > -1e-6 and m.max() < 1 + 1e-3
        # candidates listed in the same segment share one window: their masses sum to at most 1
        for pos in np.unique(f[ww > 0][:, 30]):
            same = (ww > 0) & (f[:, 30] == pos) & (f[:, 31] == f[ww > 0][f[ww > 0][:, 30] == pos][0, 31])
            ms = (np.exp(g[same].astype(np.float64)) - select_ref.LABEL_EPS) / K
            assert ms.sum() < 1 + 1e-2
        assert np.all(f[ww == 0] == 0)
    assert losses[-1] < losses[0]


def test_learned_run_selects_k(trained, small_case):
    from glimpse_amd import hip
    panel, net, _, _, _ = trained
    P, gl, flat = small_case
    prm = hip.default_params(burnin=2, main_iter=3, kpbwt=400)
    run = hip.Run(panel, gl[2:], flat[2:], prm)
    run.set_selector(net, 100, min_mass=0.0)
    run.execute()
    out = run.fetch()
    st = out["stats"]
    assert st["learned_sel"] == 2 * (prm.burnin + prm.main_iter) and st["pool_kept"] == 0
    assert np.all(out["nstates"][:, 1:] == 100)
    ds = out["gp1"] + 2 * (1 - out["gp0"] - out["gp1"])
    assert np.all(np.isfinite(ds)) and ds.min() >= -1e-4 and ds.max() <= 2 + 1e-4
    # the fallback: with min_mass 1.0 a target keeps its whole pool unless the
    # network puts no mass at all outside the kept states
    run.set_selector(net, 100, min_mass=1.0)
    run.execute()
    o2 = run.fetch()
    ns = o2["nstates"][:, 1:]
    assert set(np.unique(ns)) <= {100, 400}
    assert o2["stats"]["pool_kept"] == np.sum(ns == 400) >= 1
    assert o2["stats"]["learned_sel"] == np.sum(ns == 100)


def _collect(seed, T=10, K=400, n_haps=2000, n_sites=4000):
    from glimpse_amd import hip, synth
    P = synth.make_panel(n_haps, n_sites, T, length_mb=n_sites / 5000.0, seed=seed)
    gl, flat = synth.make_reads(P["tgt_haps"], 1.0, seed=seed + 1)
    panel = hip.Panel(P)
    run = hip.Run(panel, gl, flat, hip.default_params(burnin=3, main_iter=4, kpbwt=K))
    run.collect(1)
    run.execute()
    return panel, run, run.collected()


def test_learned_selector_gains_over_pbwt_order():
    """On a held-out panel and targets, the trained network ranks the candidates
    listed around each stored site better than PBWT order does: its first 20%
    hold more of the large-K HMM's posterior mass in that segment than the
    first 20% in PBWT order (the ones plain PBWT selection keeps at that K)."""
    from glimpse_amd import hip
    K = 400
    _, run_a, _ = _collect(101)
    net = hip.Mlp(seed=7)
    losses = net.train_run(run_a, epochs=10, batch=2048, lr=2e-3, seed=8)
    _, _, (X, y, w) = _collect(202)
    pred = net.forward(X)
    learned, pbwt = [], []
    for g in range(X.shape[0] // K):
        sl = slice(g * K, (g + 1) * K)
        keep = w[sl] > 0
        Xg, lab, sc = X[sl][keep], y[sl][keep], pred[sl][keep]
        seg = Xg[:, 30] * 1e4 + Xg[:, 31]          # candidates listed in the same segment
        for s_ in np.unique(seg):
            idx = np.nonzero(seg == s_)[0]          # in PBWT insertion order
            if len(idx) < 10:
                continue
            q = max(1, len(idx) // 5)
            learned.append(select_ref.captured_mass(lab[idx], sc[idx], q))
            pbwt.append(select_ref.captured_mass(lab[idx], np.arange(len(idx)), q))
    learned, pbwt = np.array(learned), np.array(pbwt)
    print("train loss %s; segment mass in the first 20%%: learned %.3f, PBWT order %.3f (won %d of %d)" % (
        np.round(losses, 3), learned.mean(), pbwt.mean(), np.sum(learned > pbwt), len(pbwt)))
    assert learned.mean() > pbwt.mean() + 0.02
    assert np.mean(learned >= pbwt) > 0.6
