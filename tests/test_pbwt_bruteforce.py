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

At 40-64 haplotypes every polymorphic site has a minor-allele frequency of at
least 1/64, above the default sparse_maf of 0.001, so these panels are all
common sites (the rare_frac argument only shapes the allele spectrum); rare
runs are pinned in test_pbwt_rare_bruteforce.py.
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


def _brute_lists(X, hq, cm, haps, stored, start, depth, modulo, N):
    """pbwt_states[t][o] of a panel without rare sites, from brute-force prefix
    orders: the reference's event rule (haplotype_set.cpp:445-490) and
    neighbour rule (selectK, :682-728) applied to orders obtained by sorting."""
    T = haps.shape[0] // 2
    nhq = len(hq)
    orders = [list(range(N))]                           # order before the first site, then after each
    for i in range(nhq):
        cols = hq[i::-1]
        keys = [tuple(X[h, cols]) + (h,) for h in range(N)]
        orders.append(sorted(range(N), key=lambda h: keys[h]))
    zeros = [int(np.sum(X[:, hq[i]] == 0)) for i in range(nhq)]
    lists = [[[] for _ in range(depth)] for _ in range(T)]

    def select(t, order, idx, f, g, ref, a):
        d_up, d_down = idx - f, g - idx
        if d_up < depth and d_down < depth:
            nu, nd = d_up, d_down
        elif d_up >= depth and d_down >= depth:
            nu, nd = depth, depth
        elif d_up < depth:
            nu, nd = 0, depth - d_up
        else:
            nu, nd = depth - d_down, 0
        down = []
        for j in range(nd):
            p = idx + j
            if p < N and ((p >= ref) == bool(a)):
                down.append(order[p])
            else:
                break
        up = []
        for j in range(nu):
            p = idx - 1 - j
            if p >= 0 and ((p >= ref) == bool(a)):
                up.append(order[p])
            else:
                break
        for o in range(depth):
            if o < len(down):
                lists[t][o].append(down[o])
            if o < len(up):
                lists[t][o].append(up[o])

    state = [dict(idx=int(start[h]), f=0, g=N, last=0, allele=0) for h in range(2 * T)]
    for i in range(nhq):
        k = int(hq[i])
        after = orders[i + 1]
        for h in range(2 * T):
            s = state[h]
            a = int(haps[h, k])
            span = [int(x) for x in hq if s["last"] <= x <= k]
            match = [p for p, ph in enumerate(after) if np.array_equal(X[ph, span], haps[h, span])]
            # new insertion index: the target's rank in the order after k (start breaks ties)
            tkey = tuple(haps[h, hq[i::-1]]) + (start[h] - 0.5,)
            keys = [tuple(X[ph, hq[i::-1]]) + (ph,) for ph in range(N)]
            idx_new = sum(1 for ph in range(N) if keys[ph] < tkey)
            if not match:
                # reset: neighbours in the order entering k, at the old position
                if stored[i] or cm[k - 1] - cm[s["last"]] > modulo / 2.0:
                    select(h // 2, orders[i], s["idx"], s["f"], s["g"], zeros[i - 1] if i > 0 else 0, s["allele"])
                f, g = (zeros[i], N) if a else (0, zeros[i])
                s["last"] = k
            else:
                f, g = match[0], match[-1] + 1
                assert match == list(range(f, g)), "match set not contiguous"
                if stored[i]:
                    s["idx"] = idx_new
                    select(h // 2, after, idx_new, f, g, zeros[i], a)
                    f, g = (zeros[i], N) if a else (0, zeros[i])
                    s["last"] = k
            s["idx"], s["f"], s["g"], s["allele"] = idx_new, f, g, a
    return lists


@pytest.mark.parametrize("N,L,lq,modulo,T", [(48, 160, 0, 0.05, 2), (56, 200, 6, 0.02, 2), (40, 140, 0, 0.1, 3)])
def test_selection_lists_are_brute_force(N, L, lq, modulo, T):
    """The selection lists (which panel haplotypes, at which rank, in which
    order) on panels without rare sites: the oracle's sweep against orders
    obtained by sorting, with the reference's rules for when to select (empty
    interval and a stored site or a match longer than modulo/2 cM; a stored
    site) and whom (up to depth neighbours each side of the insertion index
    that carry the target's allele)."""
    P = synth.make_panel(N, L, T, length_mb=L / 2500.0, rare_frac=0.0, flip_frac=0.05, seed=N * 5 + L,
                         keep_dense=True, sparse_maf=0.0001)
    assert P["common"].all()
    if lq:
        P["lq"] = (np.arange(P["L"]) % lq == 3).astype(np.uint8)
    rng = np.random.default_rng(N * 7 + L)
    haps = (P["tgt_haps"] ^ (rng.random(P["tgt_haps"].shape) < 0.03)).astype(np.uint8)
    OP = orc.Panel(P, sparse_maf=0.0001)
    depth = 4
    eff, stored = _stored(OP, depth, modulo, 20, rng)
    start = rng.integers(0, N, 2 * T).astype(np.int32)
    got = OP.pbwt_select(haps, stored, start, depth=depth, modulo=eff)
    hq = np.nonzero(~P["lq"].astype(bool))[0]
    cm = np.maximum(P["cm"].astype(np.float32), np.float32(0))
    want = _brute_lists(P["dense"], hq, cm, haps, stored, start, depth, eff, N)
    n = 0
    for t in range(T):
        for o in range(depth):
            assert [int(x) for x in got[t][o]] == want[t][o], "target %d rank %d" % (t, o)
            n += len(want[t][o])
    assert n > 20
