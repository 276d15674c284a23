"""Pins the oracle's phasing HMM (phasing_hmm.h:140-447, phasing_hmm.cpp:65-302)
against a dense float64 restatement of the same model on tiny cases (no GPU).

The hidden state at a phasing locus is (conditioning state k, hypothesis h),
h in 0..7 naming the alleles the modelled haplotype carries at the three hets
of its segment.  Emission: at a het of ordinal j (mod 3) ee if state k's
allele equals the allele hypothesis h carries there (1 - bit (2-j) of h), else
ed; at a peak hom 1 if the state's allele is the genotype's, else ed/ee; a flat
het has none.  Transitions between loci q and q+1 (yt from the genetic
distance, nt = 1 - yt): inside a segment h is kept and k switches LS-style,
nt*[k=k'] + yt/K; across a segment boundary staying on k frees h while
switching k keeps it, nt*[k=k'] + yt/K*[h=h'].  From the dense forward and
backward messages (each including its locus' emission) follow the quantities
the sampler uses: the per-hypothesis share of the backward message at every
segment start, the joint hypothesis probabilities across every boundary
(TRANS_HAP: alpha at the segment's last locus, any-to-any hypotheses through
nt*[k=k'] + yt/K, beta at the next segment's first locus), and P(allele 1 |
h) at every flat het.  These do not depend on the random draws.
"""
import numpy as np
import pytest

import orc
from glimpse_amd import synth


def _loci(poly, h0, h1, flat, lq):
    """Phasing loci in site order: (site, kind, het ordinal mod 3 or alt, segment)."""
    out, nhet = [], 0
    for s in poly:
        a0, a1 = int(h0[s]), int(h1[s])
        data = not flat[s] and not lq[s]
        if data and a0 != a1:
            seg = nhet // 3
            out.append((s, "het", nhet % 3, seg))
            nhet += 1
        elif data:
            out.append((s, "hom", a0, 0 if nhet == 0 else (nhet - 1) // 3))
        elif a0 != a1:
            out.append((s, "flat", a0, 0 if nhet == 0 else (nhet - 1) // 3))
    return out


def _dense_phase(alle, loci, cm, N, err_phase=1e-4, ne=100000):
    K = alle.shape[0]
    ed = np.float64(np.float32(err_phase))
    ee = np.float64(np.float32(1.0 - np.float64(np.float32(err_phase))))
    nrho = -0.04 * ne / N
    Q = len(loci)
    E = np.ones((Q, K, 8))
    for q, (s, kind, v, _) in enumerate(loci):
        a = alle[:, s].astype(int)
        if kind == "het":
            carried = np.array([1 - ((h >> (2 - v)) & 1) for h in range(8)])
            E[q] = np.where(a[:, None] == carried[None, :], ee, ed)
        elif kind == "hom":
            E[q] = np.where(a == v, 1.0, ed / ee)[:, None] * np.ones(8)[None, :]
    yt = [np.float64(np.float32(np.clip(-np.expm1(nrho * (cm[loci[q + 1][0]] - cm[loci[q][0]])), 1e-7, 1 - 1e-7)))
          for q in range(Q - 1)]
    seg = [l[3] for l in loci]

    def step(x, t, same):
        if same:   # hypothesis kept, LS switch among states
            return (1 - t) * x + t / K * x.sum(axis=0, keepdims=True)
        return (1 - t) * x.sum(axis=1, keepdims=True) + t / K * x.sum(axis=0, keepdims=True)

    A = np.zeros((Q, K, 8))
    B = np.zeros((Q, K, 8))
    A[0] = E[0] / E[0].sum()
    for q in range(1, Q):
        A[q] = E[q] * step(A[q - 1], yt[q - 1], seg[q] == seg[q - 1])
        A[q] /= A[q].sum()
    B[Q - 1] = E[Q - 1] / E[Q - 1].sum()
    for q in range(Q - 2, -1, -1):
        B[q] = E[q] * step(B[q + 1], yt[q], seg[q] == seg[q + 1])
        B[q] /= B[q].sum()
    nseg = seg[-1] + 1
    starts = [seg.index(s) for s in range(nseg)]
    seg_sums = np.array([B[q].sum(axis=0) / B[q].sum() for q in starts])
    trans = []
    for s in range(nseg - 1):
        q = starts[s + 1] - 1
        t = yt[q]
        hp = (1 - t) * A[q].T @ B[q + 1] + t / K * np.outer(A[q].sum(axis=0), B[q + 1].sum(axis=0))
        trans.append(hp / hp.sum())
    of1 = []
    for q, (s, kind, v, _) in enumerate(loci):
        if kind == "flat":
            w = A[q] * B[q]
            one = alle[:, s] == 1
            of1.append(np.clip(w[one].sum(axis=0) / w.sum(axis=0), 0.0, 1.0))
    return seg_sums, np.array(trans).reshape(-1, 8, 8), np.array(of1).reshape(-1, 8)


@pytest.mark.parametrize("K,seed", [(1, 11), (3, 12), (8, 13), (17, 14), (40, 15)])
def test_phasing_hmm_matches_dense_float64(K, seed):
    N, L = 1200, 160
    P = synth.make_panel(N, L, 1, length_mb=0.08, rare_frac=0.3, seed=seed, keep_dense=True)
    gl, flat = synth.make_reads(P["tgt_haps"], 1.0, seed=seed)
    P["lq"] = (np.arange(P["L"]) % 11 == 5).astype(np.uint8)
    OP = orc.Panel(P)
    rng = np.random.default_rng(seed)
    states = np.sort(rng.choice(N, K, replace=False))
    h0 = P["tgt_haps"][0] ^ (rng.random(L) < 0.05)
    h1 = P["tgt_haps"][1] ^ (rng.random(L) < 0.05)
    poly = OP.condset_from_states(states)["poly"]
    loci = _loci(poly, h0, h1, flat[0], P["lq"])
    assert sum(l[1] == "het" for l in loci) >= 6 and any(l[1] == "flat" for l in loci), "case too small"
    seg, hp, of1 = OP.phase_trace(states, flat[0], h0, h1)
    want_seg, want_hp, want_of1 = _dense_phase(P["dense"][states], loci, P["cm"], N)
    assert seg.shape == want_seg.shape and hp.shape == want_hp.shape and of1.shape == want_of1.shape
    assert np.abs(seg - want_seg).max() < 2e-5
    assert np.abs(hp - want_hp).max() < 2e-5
    assert np.abs(of1 - want_of1).max() < 2e-5


def test_dense_check_detects_a_wrong_boundary_rule():
    """The pin is sensitive to the segment-boundary transition: letting the
    hypothesis change on a stay inside a segment too (every step a COLLAPSE)
    moves the quantities well outside the tolerance."""
    N, L, K, seed = 1200, 160, 8, 13
    P = synth.make_panel(N, L, 1, length_mb=0.08, rare_frac=0.3, seed=seed, keep_dense=True)
    gl, flat = synth.make_reads(P["tgt_haps"], 1.0, seed=seed)
    P["lq"] = (np.arange(P["L"]) % 11 == 5).astype(np.uint8)
    OP = orc.Panel(P)
    rng = np.random.default_rng(seed)
    states = np.sort(rng.choice(N, K, replace=False))
    h0 = P["tgt_haps"][0] ^ (rng.random(L) < 0.05)
    h1 = P["tgt_haps"][1] ^ (rng.random(L) < 0.05)
    poly = OP.condset_from_states(states)["poly"]
    loci = _loci(poly, h0, h1, flat[0], P["lq"])
    seg, hp, _ = OP.phase_trace(states, flat[0], h0, h1)
    wrong = [(s, k, v, i) for i, (s, k, v, _) in enumerate(loci)]   # every locus its own segment
    w_seg, _, _ = _dense_phase(P["dense"][states], wrong, P["cm"], N)
    assert np.abs(seg[0] - w_seg[0]).max() > 1e-3
