"""GPU parity: conditioning sets (bit-exact) and the haploid forward-backward
(HP within 1e-4 absolute of the oracle for identical states and likelihoods)."""
import numpy as np
import pytest

import orc
from conftest import hvar_bytes_to_words, mask_words

pytestmark = pytest.mark.gpu

FB_TOL = 1e-4   # absolute, on haplotype posteriors


def _state_sets(N, rng):
    sets = [np.arange(N)]
    for K in (1, 7, 37, 64, 65, 300, 1000, 2000):
        K = min(K, N)
        sets.append(np.sort(rng.choice(N, K, replace=False)))
    return sets


@pytest.fixture(scope="module")
def panels(small_case):
    from glimpse_amd import hip
    P, gl, flat = small_case
    return P, gl, flat, orc.Panel(P), hip.Panel(P, with_pbwt=False)


def test_condset_bit_exact(panels):
    P, gl, flat, OP, GP = panels
    rng = np.random.default_rng(1)
    sets = _state_sets(P["N"], rng)
    got = GP.condset_batch(sets)
    for s, g in zip(sets, got):
        ref = OP.condset_from_states(s)
        K = len(s)
        assert np.array_equal(g["poly"], ref["poly"]), "polymorphic site list differs (K=%d)" % K
        assert np.array_equal(mask_words(g["hvar"], K), hvar_bytes_to_words(ref["hvar"], K)), "Hvar differs (K=%d)" % K
        # transitions: double expm1 on the device vs glibc; compare as float bits, allow 1 ulp
        d = np.abs(g["t"].view(np.int32).astype(np.int64) - ref["t"].view(np.int32).astype(np.int64))
        assert d.max(initial=0) <= 1, "transition t differs by more than 1 ulp (K=%d)" % K


def test_fb_posteriors(panels):
    P, gl, flat, OP, GP = panels
    rng = np.random.default_rng(2)
    sets = _state_sets(P["N"], rng)
    T = flat.shape[0]
    HLs, flats, refs = [], [], []
    for i, s in enumerate(sets):
        t = i % T
        other = P["tgt_haps"][2 * t + 1] if i % 2 else None
        HL = orc.hap_likelihoods(gl[t], flat[t], other)
        HLs.append(HL)
        flats.append(flat[t])
        refs.append(OP.impute_hap(s, HL, flat[t]))
    HP = GP.impute_batch(sets, np.stack(HLs), np.stack(flats))
    for i, s in enumerate(sets):
        err = np.abs(HP[i] - refs[i]).max()
        assert err <= FB_TOL, "K=%d: max |HP - oracle| = %g" % (len(s), err)
        assert np.allclose(HP[i][0::2] + HP[i][1::2], 1.0, atol=1e-5)


def test_fb_all_flat_and_lq(panels):
    """Edge cases: every site flat (no emissions) and low-quality sites."""
    P, gl, flat, OP, GP = panels
    from glimpse_amd import hip
    rng = np.random.default_rng(3)
    s = np.sort(rng.choice(P["N"], 500, replace=False))
    fl = np.ones(P["L"], np.uint8)
    HL = orc.hap_likelihoods(gl[0], fl)
    HP = GP.impute_batch([s], HL[None], fl[None])[0]
    ref = OP.impute_hap(s, HL, fl)
    assert np.abs(HP - ref).max() <= FB_TOL
    # low-quality flags: a panel copy with every 7th site LQ
    P2 = dict(P)
    P2["lq"] = (np.arange(P["L"]) % 7 == 3).astype(np.uint8)
    OP2, GP2 = orc.Panel(P2), hip.Panel(P2, with_pbwt=False)
    HL = orc.hap_likelihoods(gl[1], flat[1])
    HP = GP2.impute_batch([s], HL[None], flat[1][None])[0]
    ref = OP2.impute_hap(s, HL, flat[1])
    assert np.abs(HP - ref).max() <= FB_TOL


@pytest.mark.parametrize("shape,K", [("1x4", 200), ("2x2", 256), ("4x1", 130), ("1x8", 300), ("2x4", 500),
                                     ("4x2", 400), ("8x1", 512), ("1x16", 900), ("2x8", 600), ("4x4", 1000),
                                     ("8x2", 700), ("16x1", 1024), ("2x16", 1300), ("4x8", 2000),
                                     ("8x4", 1500), ("16x2", 1999)])
def test_fb_every_launch_shape(panels, monkeypatch, shape, K):
    """Every forward-backward launch shape (waves x words per lane) against the oracle."""
    monkeypatch.setenv("GH_FB_SHAPE", shape)
    P, gl, flat, OP, GP = panels
    rng = np.random.default_rng(K)
    s = np.sort(rng.choice(P["N"], K, replace=False))
    HL = orc.hap_likelihoods(gl[0], flat[0], P["tgt_haps"][1])
    HP = GP.impute_batch([s], HL[None], flat[0][None])[0]
    ref = OP.impute_hap(s, HL, flat[0])
    assert np.abs(HP - ref).max() <= FB_TOL


@pytest.mark.parametrize("shape", ["2x16", "4x8", "8x4", "16x2"])
def test_fb_mixed_sizes_one_launch(panels, monkeypatch, shape):
    """Items of 1..2000 states in one launch of a 2048-state shape: waves skip the
    word slots an item does not fill, and every item still matches the oracle."""
    monkeypatch.setenv("GH_FB_SHAPE", shape)
    P, gl, flat, OP, GP = panels
    rng = np.random.default_rng(11)
    sets = _state_sets(P["N"], rng)
    T = flat.shape[0]
    HLs = [orc.hap_likelihoods(gl[i % T], flat[i % T]) for i in range(len(sets))]
    HP = GP.impute_batch(sets, np.stack(HLs), np.stack([flat[i % T] for i in range(len(sets))]))
    for i, s in enumerate(sets):
        ref = OP.impute_hap(s, HLs[i], flat[i % T])
        err = np.abs(HP[i] - ref).max()
        assert err <= FB_TOL, "shape %s K=%d: max |HP - oracle| = %g" % (shape, len(s), err)


@pytest.fixture(scope="module")
def big_panels():
    """5,000 haplotypes, so conditioning sets of up to 4,096 states fit."""
    from glimpse_amd import hip, synth
    P = synth.make_panel(5000, 1500, 2, length_mb=0.5, rare_frac=0.2, flip_frac=0.05, seed=17)
    gl, flat = synth.make_reads(P["tgt_haps"], 1.0, seed=18)
    return P, gl, flat, orc.Panel(P), hip.Panel(P, with_pbwt=False)


@pytest.mark.parametrize("shape,K", [("8x8", 4096), ("16x4", 3000), ("8x8", 2049)])
def test_fb_sets_above_2048_states(big_panels, monkeypatch, shape, K):
    """Conditioning sets of up to 4,096 states (64 words): the 8x8 and 16x4 shapes."""
    monkeypatch.setenv("GH_FB_SHAPE", shape)
    P, gl, flat, OP, GP = big_panels
    rng = np.random.default_rng(K)
    sets = [np.sort(rng.choice(P["N"], K, replace=False)), np.sort(rng.choice(P["N"], 700, replace=False))]
    HLs = [orc.hap_likelihoods(gl[i], flat[i]) for i in range(2)]
    HP = GP.impute_batch(sets, np.stack(HLs), np.stack([flat[0], flat[1]]))
    for i, s in enumerate(sets):
        ref = OP.impute_hap(s, HLs[i], flat[i])
        err = np.abs(HP[i] - ref).max()
        assert err <= FB_TOL, "shape %s K=%d: max |HP - oracle| = %g" % (shape, len(s), err)
