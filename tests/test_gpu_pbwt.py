"""GPU parity: sparse-PBWT state selection is bit-exact with the oracle
(haplotype_set::matchHapsFromCompressedPBWTSmall lists, and the Kpbwt
conditioning set compactSelection takes from them)."""
import numpy as np
import pytest

import orc
from glimpse_amd import synth

pytestmark = pytest.mark.gpu

CASES = [
    # N, L, rare_frac, lq_every, modulo, T
    (1000, 3000, 0.0, 0, 0.05, 4),
    (2500, 3000, 0.5, 0, 0.1, 4),
    (2500, 2500, 0.6, 6, 0.1, 3),
    (3000, 4000, 0.3, 0, 0.02, 5),
    # 80 target haplotypes: two walk workgroups of 64 walker lanes, the second
    # one partly idle (its loaders still stage alleles for the idle lanes)
    (1500, 3000, 0.4, 0, 0.05, 40),
]


def _stored(OP, depth, modulo, kpbwt, rng):
    eff, grp = OP.groups(depth, modulo, kpbwt)
    st = np.zeros(OP.n_common_hq, np.uint8)
    lo = 0
    for g in grp:
        st[rng.integers(lo, g)] = 1
        lo = g
    return eff, st


@pytest.mark.parametrize("N,L,rare,lq,modulo,T", CASES)
def test_pbwt_select_bit_exact(N, L, rare, lq, modulo, T):
    from glimpse_amd import hip
    P = synth.make_panel(N, L, T, length_mb=L / 2500.0, rare_frac=rare, flip_frac=0.05, seed=N * 7 + L)
    if lq:
        P["lq"] = (np.arange(P["L"]) % lq == 2).astype(np.uint8)
    rng = np.random.default_rng(N + L)
    haps = (P["tgt_haps"] ^ (rng.random(P["tgt_haps"].shape) < 0.01)).astype(np.uint8)
    OP = orc.Panel(P)
    GP = hip.Panel(P, with_pbwt=True)
    depth, kpbwt = 12, 400
    eff, stored = _stored(OP, depth, modulo, kpbwt, rng)
    start = rng.integers(0, P["N"], 2 * T).astype(np.int32)
    ref = OP.pbwt_select(haps, stored, start, depth=depth, modulo=eff)
    got, states = GP.pbwt_select(haps, stored, start, depth=depth, modulo=eff, kpbwt=kpbwt)
    total = 0
    for t in range(T):
        for d in range(depth):
            assert np.array_equal(got[t][d], ref[t][d]), "target %d rank %d list differs" % (t, d)
            total += len(ref[t][d])
        want = OP.condset(ref[t], kpbwt)["states"]
        assert np.array_equal(states[t], want), "target %d compacted state set differs" % t
    assert total > 0


@pytest.mark.parametrize("every", [2, 5, 64])
@pytest.mark.parametrize("N,L,rare,lq,modulo,T", [CASES[2], CASES[3]])
def test_pbwt_select_sparse_prefix_arrays(monkeypatch, every, N, L, rare, lq, modulo, T):
    """Prefix arrays stored only every `every` HQ common sites (what large
    panels use): neighbour ids are found by stepping back through the columns
    and rare-run rearrangements, and must still be bit-exact."""
    monkeypatch.setenv("GH_PBWT_CKPT_EVERY", str(every))
    test_pbwt_select_bit_exact(N, L, rare, lq, modulo, T)
