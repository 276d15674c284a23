"""Pins the oracle's sparse-PBWT sweep for target haplotypes that carry rare
alleles, by brute force on tiny panels (no GPU): init_rare, select_rare_pd_fg
and init_common's rare-to-common branch (haplotype_set.cpp:492-680), which
the model of test_pbwt_rare_bruteforce.py (targets without rare alleles)
leaves out.

Model.  Orders are obtained by sorting: the panel haplotypes by their reversed
column sequence, ties by panel index, where a column is the allele at an HQ
common site, the small PBWT's symbol at a rare run's site (the carrier's
allele: build_sparsePBWT sets it to the major allele and flips it for the
site's SvarRef haplotypes, ref_haplotype_set.cpp:90-94), and after each run a
carrier flag (the run's carriers follow the others).  A target haplotype's
insertion index is its rank under that sort with its own columns: at a run
site 1 when it carries the site's non-major allele (select_rare_pd_fg sends
the SvarTar targets to the symbol-1 block, haplotype_set.cpp:501-505, so at
a site whose major allele is 1 such a target joins the carriers of allele 1:
the reference's indexing, kept; a target outside the run takes the major
allele, as the non-carriers hold it), and its carrier flag being whether it
carries a non-major allele in the run.  Match sets are kept as sets of panel haplotypes and mapped site by site
(the members with the target's allele or symbol); intervals are their
positions in the order, so no rank arithmetic of the oracle is reused.  The
reset rules are the reference's, stated on sets:
- a run's carriers: a carrier target keeps the carriers of its set, or all
  of them when none is in it (last reset: the run's first site); at each run
  site the members with its symbol, or when none is left every carrier with
  that symbol (last reset: that site; the block can be empty when no panel
  haplotype carries the target's rare allele, and an empty set then sits at
  the block's boundary: the small order's end for symbol 1, its start for 0);
- the run's end: a carrier target's interval is its set's positions, except
  that it starts at 0 when its last reset is at or after its last non-major
  site (the reference's rule); an empty interval selects with the order and
  interval from before the run and restarts as the whole panel (last reset:
  the site before the common site); a non-carrier keeps its set minus the
  carriers, and an empty set selects only after a match longer than modulo/5
  cM and restarts as every non-carrier;
- common sites: select_common_pd_fg's rules (test_pbwt_rare_bruteforce.py).
The oracle's (idx, f, g) after every HQ common site and its selection lists
must equal the model's.  Planted in the oracle, the run-end rule with > for
>=, or a carrier target's index one lower at the run's end, fails every case.
An interval that is empty at the run's end (its selection branch) cannot
arise from these rules: an empty set comes from an empty block at the run's
last site, whose reset is then at or after the target's last non-major site,
so the interval starts at 0 and is not empty."""
import numpy as np
import pytest

import orc
from glimpse_amd import synth

F32 = np.float32


def _panel(N, L, T, lq_every, seed):
    P = synth.make_panel(N, L, T, length_mb=L / 2500.0, rare_frac=0.3, flip_frac=0.05, seed=seed, keep_dense=True,
                         sparse_maf=0.05)
    lq = np.zeros(P["L"], np.uint8)
    if lq_every:
        lq[np.arange(P["L"]) % lq_every == 2] = 1
    common = P["common"].astype(bool)
    first = int(np.argmax(common & (lq == 0)))
    lq[:first] = 1          # the sweep starts at a common site
    P["lq"] = lq
    return P


def _model(P, haps, stored, start, depth, modulo):
    X = P["dense"].astype(np.int64)
    N, L = X.shape
    common = P["common"].astype(bool)
    lq = P["lq"].astype(bool)
    major = P["major"].astype(np.int64)
    cm = np.maximum(P["cm"].astype(np.float32), F32(0))
    hq = [k for k in range(L) if common[k] and not lq[k]]
    allc = [k for k in range(L) if common[k]]      # common2tot, LQ common sites included
    nhq = len(hq)
    runs = [[] for _ in range(nhq)]
    i = -1
    for k in range(L):
        if lq[k]:
            continue
        if common[k]:
            i += 1
        elif i >= 0 and i + 1 < nhq:
            runs[i + 1].append(k)
    SYM = X                                               # small-PBWT symbol of a panel haplotype: its allele
    NMAJ = (X != major[None, :])                          # carries the non-major allele
    TSYM = (haps.astype(np.int64) != major[None, :]).astype(np.int64)   # a target's symbol: SvarTar membership
    carriers = [set(np.nonzero(NMAJ[:, r].any(axis=1))[0].tolist()) if r else set() for r in runs]
    zeros = [int(np.sum(X[:, hq[i]] == 0)) for i in range(nhq)]
    T2 = haps.shape[0]
    T = T2 // 2
    pcols = [[] for _ in range(N)]
    tcols = [[] for _ in range(T2)]

    def order_now():
        return sorted(range(N), key=lambda p: (tuple(reversed(pcols[p])), p))

    def tpos(h, order):
        key = (tuple(reversed(tcols[h])), start[h] - 0.5)
        return sum(1 for p in order if (tuple(reversed(pcols[p])), p) < key)

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
        down, up = [], []
        for j in range(nd):
            p = idx + j
            if p < N and ((p >= ref) == bool(a)):
                down.append(order[p])
            else:
                break
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

    def interval(order, M):
        pos = [p for p, ph in enumerate(order) if ph in M]
        assert pos, "empty match set"
        assert pos == list(range(pos[0], pos[-1] + 1)), "match set not contiguous in the order"
        return pos[0], pos[-1] + 1

    st = [dict(idx=int(start[h]), M=set(range(N)), last=0, last_rare=0, allele=0) for h in range(T2)]
    stats = dict(carrier_runs=0, start_at_zero=0, empty_block=0, rare_end_select=0)
    order = list(range(N))
    trace = []
    for i in range(nhq):
        k = hq[i]
        if i > 0 and runs[i]:
            prev = order
            C = carriers[i]
            ns = len(C)
            inr = [bool(TSYM[h, runs[i]].any()) for h in range(T2)]
            for c in runs[i]:
                for p in range(N):
                    pcols[p].append(int(SYM[p, c]))
                for h in range(T2):   # a target outside the run ties with the non-carriers (the major allele)
                    tcols[h].append(int(TSYM[h, c]) if inr[h] else int(major[c]))
            for p in range(N):
                pcols[p].append(int(p in C))
            for h in range(T2):
                tcols[h].append(int(inr[h]))
            entering = order_now()
            for h in range(T2):
                s = st[h]
                if not inr[h]:
                    M = s["M"] - C
                    if not M:
                        ps = allc[i - 1]
                        if cm[ps] - cm[s["last"]] > F32(modulo) / F32(5.0):
                            f0, g0 = interval(prev, s["M"])
                            select(h // 2, prev, s["idx"], f0, g0, zeros[i - 1], s["allele"])
                        M = set(range(N)) - C
                        s["last"] = ps
                else:
                    stats["carrier_runs"] += 1
                    Ms = s["M"] & C
                    if not Ms:
                        Ms = set(C)
                        s["last"] = runs[i][0]
                    last_sym = 0
                    for c in runs[i]:
                        a = int(TSYM[h, c])
                        if a:
                            s["last_rare"] = c
                        nxt = {p for p in Ms if SYM[p, c] == a}
                        if not nxt:
                            nxt = {p for p in C if SYM[p, c] == a}
                            s["last"] = c
                        Ms = nxt
                        last_sym = a
                    if Ms:
                        lo, hi = interval(entering, Ms)
                    else:
                        stats["empty_block"] += 1
                        lo = hi = N if last_sym else N - ns
                    f = lo
                    if s["last"] >= s["last_rare"]:
                        f = 0
                        stats["start_at_zero"] += 1
                    g = hi
                    if g <= f:
                        stats["rare_end_select"] += 1
                        f0, g0 = interval(prev, s["M"])
                        select(h // 2, prev, s["idx"], f0, g0, zeros[i - 1], s["allele"])
                        M = set(range(N))
                        s["last"] = k - 1
                    else:
                        M = set(entering[f:g])
                s["M"] = M
                s["idx"] = tpos(h, entering)
            order = entering
        entering = order
        for p in range(N):
            pcols[p].append(int(X[p, k]))
        for h in range(T2):
            tcols[h].append(int(haps[h, k]))
        after = order_now()
        for h in range(T2):
            s = st[h]
            a = int(haps[h, k])
            idx_new = tpos(h, after)
            M = {p for p in s["M"] if X[p, k] == a}
            if not M:
                if stored[i] or cm[k - 1] - cm[s["last"]] > F32(modulo) / F32(2.0):
                    f0, g0 = interval(entering, s["M"])
                    select(h // 2, entering, s["idx"], f0, g0, zeros[i - 1] if i > 0 else 0, s["allele"])
                M = {p for p in range(N) if X[p, k] == a}
                s["last"] = k
            elif stored[i]:
                f0, g0 = interval(after, M)
                select(h // 2, after, idx_new, f0, g0, zeros[i], a)
                M = {p for p in range(N) if X[p, k] == a}
                s["last"] = k
            s["M"], s["idx"], s["allele"] = M, idx_new, a
            f, g = interval(after, M)
            trace.append((k, h, idx_new, f, g))
        order = after
    return lists, trace, stats


CASES = [(64, 240, 0, 0.02, 3, 11), (56, 220, 9, 0.05, 3, 12), (64, 260, 0, 0.1, 3, 13), (48, 200, 7, 0.02, 3, 14)]


@pytest.mark.parametrize("N,L,lq_every,modulo,T,seed", CASES)
def test_carrier_targets_positions_and_selection_are_brute_force(N, L, lq_every, modulo, T, seed):
    P = _panel(N, L, T, lq_every, seed)
    rng = np.random.default_rng(seed)
    haps = (P["tgt_haps"] ^ (rng.random(P["tgt_haps"].shape) < 0.03)).astype(np.uint8)
    OP = orc.Panel(P, sparse_maf=0.05)
    depth = 4
    eff, grp = OP.groups(depth, modulo, 16)
    stored = np.zeros(OP.n_common_hq, np.uint8)
    lo = 0
    for g in grp:
        stored[rng.integers(lo, g)] = 1
        lo = g
    start = rng.integers(0, N, 2 * T).astype(np.int32)
    want, trace, stats = _model(P, haps, stored, start, depth, eff)
    assert stats["carrier_runs"] > 20, stats
    rows = OP.pbwt_trace(haps, stored, start, depth=depth, modulo=eff)
    by = {(int(r[0]), int(r[1])): r for r in rows}
    assert len(by) == len(trace)
    for k, h, idx, f, g in trace:
        r = by[(k, h)]
        assert (int(r[2]), int(r[3]), int(r[4])) == (idx, f, g), \
            "site %d hap %d: oracle (idx, f, g) = %s, model %s" % (k, h, tuple(int(x) for x in r[2:5]), (idx, f, g))
    got = OP.pbwt_select(haps, stored, start, depth=depth, modulo=eff)
    n = 0
    for t in range(T):
        for o in range(depth):
            assert [int(x) for x in got[t][o]] == want[t][o], "target %d rank %d" % (t, o)
            n += len(want[t][o])
    assert n > 20
    print("[rare carriers] N=%d L=%d: %s, %d selected ids" % (N, L, stats, n))


def test_carrier_cases_cover_the_run_end_rules():
    """Over the cases above the model passes through every branch it states:
    the interval starting at 0, empty blocks and run-end selections."""
    tot = dict(carrier_runs=0, start_at_zero=0, empty_block=0, rare_end_select=0)
    for N, L, lq_every, modulo, T, seed in CASES:
        P = _panel(N, L, T, lq_every, seed)
        rng = np.random.default_rng(seed)
        haps = (P["tgt_haps"] ^ (rng.random(P["tgt_haps"].shape) < 0.03)).astype(np.uint8)
        OP = orc.Panel(P, sparse_maf=0.05)
        eff, grp = OP.groups(4, modulo, 16)
        stored = np.zeros(OP.n_common_hq, np.uint8)
        lo = 0
        for g in grp:
            stored[rng.integers(lo, g)] = 1
            lo = g
        start = rng.integers(0, N, 2 * T).astype(np.int32)
        _, _, stats = _model(P, haps, stored, start, 4, eff)
        for key in tot:
            tot[key] += stats[key]
    print("[rare carriers] branches:", tot)
    assert tot["carrier_runs"] > 100 and tot["start_at_zero"] > 0 and tot["empty_block"] > 0, tot
