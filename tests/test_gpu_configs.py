"""The workload configurations at their full sizes, on the GPU, checked
against the oracle (bit-exact state sets, forward-backward within 1e-4,
Gibbs runs draw for draw) and through size-independent properties.

  C1  1,000 haplotypes x 10,000 SNPs, 1 target at 1x, Kpbwt 500, 1+5+15 iterations
  C2  one full-size chunk of the batched workload: 10,000 haplotypes, 60,000
      sites, 200 targets at U(0.1, 2)x, Kpbwt U[500, 2000]
  C3  100,000 haplotypes x 150,000 sites, about 80% of sites at MAF < 0.1%,
      a chunk of more than 21 cM so PBWT selection runs every 0.5 cM, K 1,000
Each test prints the figures it checks."""
import time

import numpy as np
import pytest

import orc
from conftest import dosage_r2, nrc
from glimpse_amd import synth

pytestmark = pytest.mark.gpu


def _stored(OP, depth, modulo, kpbwt, rng):
    eff, grp = OP.groups(depth, modulo, kpbwt)
    st = np.zeros(OP.n_common_hq, np.uint8)
    lo = 0
    for g in grp:
        st[rng.integers(lo, g)] = 1
        lo = g
    return eff, st


@pytest.fixture(scope="module")
def c3():
    """C3 panel: 100,000 haplotypes x 150,000 sites over about 22 cM, 75% of
    candidate sites injected rare (the 1/x spectrum adds more), 4 targets at 1x."""
    t0 = time.time()
    P = synth.make_panel(100000, 150000, 4, length_mb=36.0, rare_frac=0.75, seed=303, fast=True)
    gl, flat = synth.make_reads(P["tgt_haps"], 1.0, seed=304, threads=0)
    print("[c3] panel %.1f s: %d sites over %.2f cM, common %.3f, MAF < 0.1%%: %.3f" % (
        time.time() - t0, P["L"], P["cm"][-1], P["common"].mean(),
        np.mean(np.minimum(P["cref"], P["calt"]) / P["N"] < 0.001)), flush=True)
    return P, gl, flat


@pytest.mark.timeout(900)
def test_c3_full_size_selection_and_posteriors(c3):
    from glimpse_amd import hip
    P, gl, flat = c3
    t0 = time.time()
    OP = orc.Panel(P)
    t1 = time.time()
    GP = hip.Panel(P, with_pbwt=True)
    t2 = time.time()
    info = GP.info()
    print("[c3] oracle panel %.1f s, GPU panel %.1f s (%.2f GB on the device)" % (
        t1 - t0, t2 - t1, info["device_bytes"] / 1e9), flush=True)
    assert P["cm"][-1] >= 21.0, "chunk shorter than 21 cM"
    assert np.mean(np.minimum(P["cref"], P["calt"]) / P["N"] < 0.001) >= 0.75
    depth, kpbwt = 12, 1000
    rng = np.random.default_rng(305)
    eff, stored = _stored(OP, depth, 0.5, kpbwt, rng)
    assert eff == pytest.approx(0.5), "PBWT modulo adjusted to %g" % eff
    haps = (P["tgt_haps"] ^ (rng.random(P["tgt_haps"].shape) < 0.01)).astype(np.uint8)
    start = rng.integers(0, P["N"], haps.shape[0]).astype(np.int32)
    t0 = time.time()
    ref = OP.pbwt_select(haps, stored, start, depth=depth, modulo=eff)
    t1 = time.time()
    got, states = GP.pbwt_select(haps, stored, start, depth=depth, modulo=eff, kpbwt=kpbwt)
    t2 = time.time()
    T = haps.shape[0] // 2
    sets = []
    for t in range(T):
        for d in range(depth):
            assert np.array_equal(got[t][d], ref[t][d]), "target %d rank %d list differs" % (t, d)
        want = OP.condset(ref[t], kpbwt)["states"]
        assert np.array_equal(states[t], want), "target %d compacted state set differs" % t
        assert len(want) == kpbwt
        sets.append(want)
    print("[c3] PBWT selection: oracle %.1f s, GPU call %.2f s; %d targets bit-exact" % (t1 - t0, t2 - t1, T),
          flush=True)
    # forward-backward of each target's two haplotypes on its selected set
    HLs, fl, items = [], [], []
    for t in range(T):
        for j in range(2):
            other = P["tgt_haps"][2 * t + 1 - j] if j else None
            HLs.append(orc.hap_likelihoods(gl[t], flat[t], other))
            fl.append(flat[t])
            items.append(sets[t])
    HP = GP.impute_batch(items, np.stack(HLs), np.stack(fl))
    worst = 0.0
    for i, s in enumerate(items):
        r = OP.impute_hap(s, HLs[i], fl[i])
        worst = max(worst, float(np.abs(HP[i] - r).max()))
    print("[c3] forward-backward at K=%d on %d haplotypes: max |HP - oracle| = %.2e" % (kpbwt, len(items), worst))
    assert worst <= 1e-4
