"""GPU parity: phasing HMM (phasing_hmm::rephaseHaplotypes) with counter-mode
draws.  Given identical states, haplotypes, flat flags and draw keys, the
rephased haplotypes must equal the oracle's (the probabilities differ only in
summation order, far below the spacing of the uniform draws)."""
import numpy as np
import pytest

import orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def panels(small_case):
    from glimpse_amd import hip
    P, gl, flat = small_case
    return P, gl, flat, orc.Panel(P), hip.Panel(P, with_pbwt=False)


def _perturb(h, rate, rng):
    return (h ^ (rng.random(h.shape) < rate)).astype(np.uint8)


@pytest.mark.parametrize("K", [5, 64, 300, 1500, 2000])
def test_rephase_matches_oracle(panels, K):
    P, gl, flat, OP, GP = panels
    rng = np.random.default_rng(K)
    T = flat.shape[0]
    sets, h0s, h1s, tg = [], [], [], []
    for t in range(T):
        sets.append(np.sort(rng.choice(P["N"], min(K, P["N"]), replace=False)))
        h0s.append(_perturb(P["tgt_haps"][2 * t], 0.02, rng))
        h1s.append(_perturb(P["tgt_haps"][2 * t + 1], 0.02, rng))
        tg.append(100 + t)
    seed, it = 777, 3
    g0, g1 = GP.rephase_batch(sets, flat, np.stack(h0s), np.stack(h1s), seed, it, tg)
    for t in range(T):
        r0, r1 = OP.rephase(sets[t], flat[t], h0s[t], h1s[t], seed, it, tg[t])
        n0 = int((g0[t] != r0).sum())
        n1 = int((g1[t] != r1).sum())
        assert n0 == 0 and n1 == 0, "target %d K=%d: %d/%d haplotype alleles differ" % (t, K, n0, n1)
        # the genotype (h0 + h1) is never changed by rephasing except at flat
        # (or low-quality) hets, where phasing_hmm.cpp:269-285 resamples H0 only
        assert np.array_equal((g0[t] ^ g1[t]) | flat[t], (h0s[t] ^ h1s[t]) | flat[t])


def test_rephase_with_lq_sites(small_case):
    """Low-quality panel sites join the flat hets: no
    emission, H0 resampled (phasing_hmm.cpp:75-94).  A panel copy with every 5th site LQ."""
    from glimpse_amd import hip
    P, gl, flat = small_case
    P2 = dict(P)
    P2["lq"] = (np.arange(P["L"]) % 5 == 2).astype(np.uint8)
    OP2, GP2 = orc.Panel(P2), hip.Panel(P2, with_pbwt=False)
    rng = np.random.default_rng(5)
    T = flat.shape[0]
    for K in (150, 1200):
        sets = [np.sort(rng.choice(P["N"], K, replace=False)) for _ in range(T)]
        h0s = np.stack([_perturb(P["tgt_haps"][2 * t], 0.02, rng) for t in range(T)])
        h1s = np.stack([_perturb(P["tgt_haps"][2 * t + 1], 0.02, rng) for t in range(T)])
        g0, g1 = GP2.rephase_batch(sets, flat, h0s, h1s, 31, 4, list(range(T)))
        free = flat | P2["lq"]
        for t in range(T):
            r0, r1 = OP2.rephase(sets[t], flat[t], h0s[t], h1s[t], 31, 4, t)
            assert np.array_equal(g0[t], r0) and np.array_equal(g1[t], r1), "K=%d target %d" % (K, t)
            assert np.array_equal((g0[t] ^ g1[t]) | free[t], (h0s[t] ^ h1s[t]) | free[t])


# every launch shape the phasing kernel is built for, at a set size of its class
_SHAPES = [(1, 1, 60), (1, 2, 100), (1, 4, 250), (2, 1, 100), (2, 2, 250), (4, 1, 256), (1, 8, 450), (2, 4, 500),
           (4, 2, 400), (8, 1, 512), (2, 8, 900), (4, 4, 1000), (8, 2, 700), (16, 1, 1024), (4, 8, 2000), (8, 4, 1500),
           (16, 2, 1999)]


@pytest.mark.parametrize("nw,kpt,K", _SHAPES)
def test_rephase_every_launch_shape(panels, monkeypatch, nw, kpt, K):
    monkeypatch.setenv("GH_PHASE_SHAPE", "%dx%d" % (nw, kpt))
    P, gl, flat, OP, GP = panels
    rng = np.random.default_rng(K + nw)
    T = 2
    sets = [np.sort(rng.choice(P["N"], K, replace=False)) for _ in range(T)]
    h0s = np.stack([_perturb(P["tgt_haps"][2 * t], 0.02, rng) for t in range(T)])
    h1s = np.stack([_perturb(P["tgt_haps"][2 * t + 1], 0.02, rng) for t in range(T)])
    g0, g1 = GP.rephase_batch(sets, flat[:T], h0s, h1s, 99, 2, [7, 8])
    for t in range(T):
        r0, r1 = OP.rephase(sets[t], flat[t], h0s[t], h1s[t], 99, 2, 7 + t)
        assert np.array_equal(g0[t], r0) and np.array_equal(g1[t], r1), "shape %dx%d K=%d target %d" % (nw, kpt, K, t)


@pytest.mark.parametrize("K", [2500, 4096])
def test_rephase_sets_above_2048_states(K):
    """Phasing with conditioning sets of up to 4,096 states (the 8 x 8 shape)."""
    from glimpse_amd import hip, synth
    P = synth.make_panel(5000, 1500, 2, length_mb=0.5, rare_frac=0.2, flip_frac=0.05, seed=17)
    gl, flat = synth.make_reads(P["tgt_haps"], 1.0, seed=18)
    OP, GP = orc.Panel(P), hip.Panel(P, with_pbwt=False)
    rng = np.random.default_rng(K)
    T = 2
    sets = [np.sort(rng.choice(P["N"], K, replace=False)) for _ in range(T)]
    h0s = np.stack([_perturb(P["tgt_haps"][2 * t], 0.02, rng) for t in range(T)])
    h1s = np.stack([_perturb(P["tgt_haps"][2 * t + 1], 0.02, rng) for t in range(T)])
    g0, g1 = GP.rephase_batch(sets, flat, h0s, h1s, 5, 1, [0, 1])
    for t in range(T):
        r0, r1 = OP.rephase(sets[t], flat[t], h0s[t], h1s[t], 5, 1, t)
        assert np.array_equal(g0[t], r0) and np.array_equal(g1[t], r1), "K=%d target %d" % (K, t)
