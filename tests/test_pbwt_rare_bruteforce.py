"""Pins the oracle's sparse-PBWT sweep across rare-variant runs by brute force
on tiny panels (no GPU).  With sparse_maf 0.05 and 64 haplotypes about two
thirds of the sites are rare (minor-allele count <= 3), so the sweep spends
most of its sites in rare runs (haplotype_set.cpp:492-680: init_rare,
select_rare_pd_fg, init_common).

Model.  A rare run moves the haplotypes that carry a non-major allele at one
of its sites (the run's carriers, ref_haplotype_set.cpp:47-79) into a small
PBWT over the run and appends them, in its order, after the non-carriers at
the run's end.  That is a standard PBWT (Durbin 2014) over the run's sites
with the actual alleles followed by one virtual column, the carrier flag: the
order after any point of the sweep is the sort by reversed column sequence,
ties by panel index.  The targets here carry no rare allele (every run ends
for them in init_common's common-to-common branch), so after a run a target's
match interval holds the panel haplotypes that equal it at every HQ common
site since its last reset and carry no rare allele of the runs since then.
At a run's end a target whose interval became empty selects when the match
is longer than modulo/5 cM (haplotype_set.cpp:611-627): the neighbours of its
position in the order before the run, as selectK picks them; its interval
restarts as all non-carriers and its last reset is the previous common site
(common2tot[l - 1] with the HQ-common index l, which with LQ common sites is
an earlier common site: the reference's indexing, kept).
At common sites the rules are select_common_pd_fg's (:445-490), with the
order entering the site (after the run's rearrangement) for a reset's
selection and the previous common site's zero count as selectK's allele
split (what the reference passes).
"""
import numpy as np
import pytest

import orc
from glimpse_amd import synth


def _panel(N, L, T, lq_every, seed):
    P = synth.make_panel(N, L, T, length_mb=L / 2500.0, rare_frac=0.3, flip_frac=0.05, seed=seed, keep_dense=True,
                         sparse_maf=0.05)
    lq = np.zeros(P["L"], np.uint8)
    if lq_every:
        lq[np.arange(P["L"]) % lq_every == 2] = 1
    common = P["common"].astype(bool)
    first = int(np.argmax(common & (lq == 0)))
    lq[:first] = 1          # the sweep starts at a common site (a leading run has no common site before it)
    P["lq"] = lq
    return P


def _brute(P, haps, stored, start, depth, modulo):
    """Selection lists and the per-site states (idx, match set) of the model."""
    X = P["dense"]
    N, L = X.shape
    common = P["common"].astype(bool)
    lq = P["lq"].astype(bool)
    major = P["major"].astype(np.uint8)
    cm = np.maximum(P["cm"].astype(np.float32), np.float32(0))
    hq = [k for k in range(L) if common[k] and not lq[k]]
    allc = [k for k in range(L) if common[k]]   # every common site, LQ ones included (common2tot)
    nhq = len(hq)
    # the run before each HQ common site (non-LQ rare sites since the previous one)
    runs = [[] for _ in range(nhq)]
    i = -1
    for k in range(L):
        if lq[k]:
            continue
        if common[k]:
            i += 1
        elif i >= 0 and i + 1 < nhq:
            runs[i + 1].append(k)
    carrier = [np.zeros(N, bool) for _ in range(nhq)]
    for i in range(nhq):
        for k in runs[i]:
            carrier[i] |= X[:, k] != major[k]
    zeros = [int(np.sum(X[:, hq[i]] == 0)) for i in range(nhq)]
    T2 = haps.shape[0]

    # column sequences (panel and targets) and the orders after each event
    pcols = [[] for _ in range(N)]
    tcols = [[] for _ in range(T2)]

    def order_now():
        return sorted(range(N), key=lambda p: (tuple(reversed(pcols[p])), p))

    def tpos(h, order):
        key = (tuple(reversed(tcols[h])), start[h] - 0.5)
        return sum(1 for p in order if (tuple(reversed(pcols[p])), p) < key)

    T = T2 // 2
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

    def interval(order, members):
        pos = [p for p, ph in enumerate(order) if ph in members]
        if not pos:
            return None
        assert pos == list(range(pos[0], pos[-1] + 1)), "match set not contiguous in the order"
        return pos[0], pos[-1] + 1

    def matches(h, span_lo, runs_lo, i_hi, runs_hi):
        m = np.ones(N, bool)
        for j in range(span_lo, i_hi + 1):
            m &= X[:, hq[j]] == haps[h, hq[j]]
        for j in range(max(runs_lo, 0), runs_hi + 1):
            m &= ~carrier[j]
        return set(np.nonzero(m)[0].tolist())

    st = [dict(idx=int(start[h]), f=0, g=N, last=0, allele=0, span_lo=0, runs_lo=0) for h in range(T2)]
    order = list(range(N))
    trace = []
    for i in range(nhq):
        k = hq[i]
        if i > 0 and runs[i]:
            prev_order = order
            for c in runs[i]:
                for p in range(N):
                    pcols[p].append(int(X[p, c]))
                for h in range(T2):
                    tcols[h].append(int(haps[h, c]))
            for p in range(N):
                pcols[p].append(int(carrier[i][p]))
            for h in range(T2):
                tcols[h].append(0)
            entering = order_now()
            for h in range(T2):
                s = st[h]
                m = matches(h, s["span_lo"], s["runs_lo"], i - 1, i)
                if not m:
                    # the reference reads common2tot[l - 1] with the HQ-common index l: with LQ common
                    # sites before it, that names an earlier common site than the previous HQ one
                    prev_site = allc[i - 1]
                    if cm[prev_site] - cm[s["last"]] > np.float32(modulo) / np.float32(5.0):
                        select(h // 2, prev_order, s["idx"], s["f"], s["g"], zeros[i - 1], s["allele"])
                    s["f"], s["g"] = 0, int(N - carrier[i].sum())
                    s["last"], s["span_lo"], s["runs_lo"] = prev_site, i, i
                else:
                    s["f"], s["g"] = interval(entering, m)
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
            m = matches(h, s["span_lo"], s["runs_lo"], i, i)
            if not m:
                if stored[i] or cm[k - 1] - cm[s["last"]] > np.float32(modulo) / np.float32(2.0):
                    select(h // 2, entering, s["idx"], s["f"], s["g"], zeros[i - 1] if i > 0 else 0, s["allele"])
                s["f"], s["g"] = (zeros[i], N) if a else (0, zeros[i])
                s["last"], s["span_lo"], s["runs_lo"] = k, i, i + 1
            else:
                s["f"], s["g"] = interval(after, m)
                if stored[i]:
                    select(h // 2, after, idx_new, s["f"], s["g"], zeros[i], a)
                    s["f"], s["g"] = (zeros[i], N) if a else (0, zeros[i])
                    s["last"], s["span_lo"], s["runs_lo"] = k, i, i + 1
            s["idx"], s["allele"] = idx_new, a
            trace.append((k, h, s["idx"], s["f"], s["g"]))
        order = after
    return lists, trace, sum(len(r) for r in runs)


@pytest.mark.parametrize("N,L,lq_every,modulo,T,seed", [(64, 220, 0, 0.02, 2, 3), (56, 200, 9, 0.05, 2, 4),
                                                        (64, 240, 0, 0.1, 3, 5)])
def test_rare_runs_positions_and_selection_are_brute_force(N, L, lq_every, modulo, T, seed):
    P = _panel(N, L, T, lq_every, seed)
    common = P["common"].astype(bool)
    assert (~common).mean() > 0.4, "the panel needs rare runs"
    rng = np.random.default_rng(seed)
    haps = (P["tgt_haps"] ^ (rng.random(P["tgt_haps"].shape) < 0.03)).astype(np.uint8)
    rare = ~common
    haps[:, rare] = P["major"].astype(np.uint8)[rare]          # targets carry no rare allele
    OP = orc.Panel(P, sparse_maf=0.05)
    depth = 4
    eff, grp = OP.groups(depth, modulo, 16)
    stored = np.zeros(OP.n_common_hq, np.uint8)
    lo = 0
    for g in grp:
        stored[rng.integers(lo, g)] = 1
        lo = g
    start = rng.integers(0, N, 2 * T).astype(np.int32)
    want, trace, n_rare = _brute(P, haps, stored, start, depth, eff)
    assert n_rare > 50
    # positions and intervals after every HQ common site
    rows = OP.pbwt_trace(haps, stored, start, depth=depth, modulo=eff)
    by = {(int(r[0]), int(r[1])): r for r in rows}
    assert len(by) == len(trace)
    for k, h, idx, f, g in trace:
        r = by[(k, h)]
        assert (int(r[2]), int(r[3]), int(r[4])) == (idx, f, g), \
            "site %d hap %d: oracle (idx, f, g) = %s, model %s" % (k, h, tuple(int(x) for x in r[2:5]), (idx, f, g))
    # the selection lists, run-end selections included
    got = OP.pbwt_select(haps, stored, start, depth=depth, modulo=eff)
    n = 0
    for t in range(T):
        for o in range(depth):
            assert [int(x) for x in got[t][o]] == want[t][o], "target %d rank %d" % (t, o)
            n += len(want[t][o])
    assert n > 20
