"""Pins the oracle's sparse-PBWT sweep (haplotype_set.cpp:317-560,
matchHapsFromCompressedPBWTSmall / select_common_pd_fg) by brute force on
tiny panels (no GPU), following Durbin (2014).

Over the HQ common sites (LQ and rare sites do not move the common
interval), after site k:
* the target haplotype's insertion index idx is its rank among the panel
  haplotypes sorted by reversed prefix (alleles at k, k-1, ... back to the
  first HQ common site, ties broken by panel index), the target's own tie
  position being just before the panel haplotype its random start names;
* the match interval [f, g) of the prefix array holds exactly the panel
  haplotypes whose alleles equal the target's at every HQ common site from its
  last reset (an empty interval or a stored site) to k.
The selection lists are then the reference's rule applied to these
positions; the GPU reproduces them bit for bit (test_gpu_pbwt.py).
"""
import numpy as np
import pytest

import orc
from glimpse_amd import synth


def _stored(OP, depth, modulo, kpbwt, rng):
    eff, grp = OP.groups(depth, modulo, kpbwt)
    st = np.zeros(OP.n_common_hq, np.uint8)
    lo = 0
    for g in grp:
        st[rng.integers(lo, g)] = 1
        lo = g
    return eff, st


@pytest.mark.parametrize("N,L,rare,lq,modulo,T", [(48, 150, 0.0, 0, 0.05, 2), (64, 200, 0.4, 7, 0.02, 3),
                                                  (40, 180, 0.3, 0, 0.1, 3)])
def test_sweep_positions_and_intervals_are_brute_force(N, L, rare, lq, modulo, T):
    P = synth.make_panel(N, L, T, length_mb=L / 2500.0, rare_frac=rare, flip_frac=0.05, seed=N * 3 + L,
                         keep_dense=True)
    if lq:
        P["lq"] = (np.arange(P["L"]) % lq == 2).astype(np.uint8)
    rng = np.random.default_rng(N + L)
    haps = (P["tgt_haps"] ^ (rng.random(P["tgt_haps"].shape) < 0.02)).astype(np.uint8)
    OP = orc.Panel(P)
    eff, stored = _stored(OP, 12, modulo, 40, rng)
    start = rng.integers(0, N, 2 * T).astype(np.int32)
    rows = OP.pbwt_trace(haps, stored, start, depth=12, modulo=eff)
    X = P["dense"]                                     # [N][L]
    hq = np.nonzero(P["common"].astype(bool) & ~P["lq"].astype(bool))[0]
    pos = {int(k): i for i, k in enumerate(hq)}
    assert len(rows) == len(hq) * 2 * T
    assert (rows[:, 5] >= 0).any(), "no reset exercised"
    order_at, keys_at = {}, {}
    n_checked = 0
    for k, htr, idx, f, g, rs in rows:
        i = pos[int(k)]
        if i not in order_at:
            cols = hq[i::-1]
            keys = [tuple(X[h, cols]) + (h,) for h in range(N)]
            order_at[i] = sorted(range(N), key=lambda h: keys[h])
            keys_at[i] = keys
        order, keys = order_at[i], keys_at[i]
        tkey = tuple(haps[htr, hq[i::-1]]) + (start[htr] - 0.5,)
        assert idx == sum(1 for h in range(N) if keys[h] < tkey), "idx at site %d hap %d" % (k, htr)
        lo = pos[int(rs)] if rs >= 0 else 0
        span = hq[lo:i + 1]
        match = {h for h in range(N) if np.array_equal(X[h, span], haps[htr, span])}
        assert 0 <= f <= g <= N and set(order[f:g]) == match, "interval at site %d hap %d" % (k, htr)
        n_checked += 1
    assert n_checked > 100
