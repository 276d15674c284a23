"""CPU restatement of the learned selector's features and top-k step.

TEST INFRASTRUCTURE ONLY: imported by tests/ (and nothing else), never by the
product package.  The learned selector has no counterpart in the reference;
this file states its definition in plain numpy, site by site, so the device
kernels (bit tricks on transposed 64-bit panel words, ballots, bitonic sorts)
can be checked against it.

A candidate's pool code is id | rank << 20 | (listed by H1's walk) << 26 |
(upper neighbour) << 27.  Common sites are the panel's common rows in site
order; "words" are groups of 64 consecutive common sites (the last one
partial).  Float features are computed with the same float32 operations as the
kernels, so they agree bit for bit.
"""
import numpy as np

FEAT = 32
WIN = 8
CONF = 10          # phred gap of a confident genotype call
LABEL_EPS = 1e-3

f32 = np.float32


def _common_alleles(P):
    """[Lc][N] alleles of the panel at its common sites."""
    words = P["common_words"]
    N = P["N"]
    bits = np.unpackbits(words.view(np.uint8).reshape(words.shape[0], -1), axis=1, bitorder="little")
    return bits[:, :N].astype(bool)


def _longest_true_run(b):
    best = cur = 0
    for x in b:
        cur = cur + 1 if x else 0
        best = max(best, cur)
    return best


NEAR, FAR, RUN = 4, 16, 64    # words of 64 common sites


def _word_counts(bits, Wc):
    """Per-word count of set bits (the last word may be partial)."""
    return np.bincount(np.arange(len(bits)) // 64, weights=bits, minlength=Wc)


def features(P, H, gl, flat, pool, pool_site, K, depth):
    """X [T][kstride][32] for the pools (rows >= K[t] stay zero)."""
    T, ks = pool.shape
    L = P["L"]
    common = P["common"].astype(bool)
    c2t = np.nonzero(common)[0]
    crow = np.full(L, -1)
    crow[c2t] = np.arange(len(c2t))
    Lc = len(c2t)
    Wc = (Lc + 63) // 64
    A = _common_alleles(P)
    major = P["major"].astype(np.uint8)
    rp, ri = P["rare_ptr"], P["rare_idx"]
    X = np.zeros((T, ks, FEAT), np.float32)

    def wsites(lo, hi):
        lo, hi = max(lo, 0), min(hi, Wc)
        return f32(max(min(64 * hi, Lc) - 64 * lo, 0))

    def rng(cnt, lo, hi):
        lo, hi = max(lo, 0), min(hi, Wc)
        return f32(cnt[lo:hi].sum()) if hi > lo else f32(0)

    for t in range(T):
        h0 = H[2 * t][c2t].astype(bool)
        h1 = H[2 * t + 1][c2t].astype(bool)
        fl = flat[t][c2t].astype(bool)
        g = gl[t][c2t].astype(np.int32)
        D = ~fl
        RR = D & (g[:, 0] == 0) & (g[:, 1] >= CONF) & (g[:, 2] >= CONF)
        RA = D & (g[:, 1] == 0) & (g[:, 0] >= CONF) & (g[:, 2] >= CONF)
        AA = D & (g[:, 2] == 0) & (g[:, 0] >= CONF) & (g[:, 1] >= CONF)
        inv_d = f32(1.0) / f32(max(D.sum(), 1))
        inv_lc = f32(1.0) / f32(max(Lc, 1))
        inv_wc = f32(1.0) / f32(max(Wc, 1))
        wD, wRR, wAA, wRA = (_word_counts(m, Wc) for m in (D, RR, AA, RA))
        share0, share1 = set(), set()
        for l in range(L):
            if common[l]:
                continue
            c = ri[rp[l]:rp[l + 1]]
            if H[2 * t][l] != major[l]:
                share0.update(c.tolist())
            if H[2 * t + 1][l] != major[l]:
                share1.update(c.tolist())
        k = int(K[t])
        inv_k = f32(1.0) / f32(max(k, 1))
        for j in range(k):
            v = int(pool[t, j])
            cid = v & 0xFFFFF
            cw = max(int(crow[pool_site[t, j]]), 0) >> 6
            R = A[:, cid]
            x0, x1 = R ^ h0, R ^ h1
            w0, w1 = _word_counts(x0, Wc), _word_counts(x1, Wc)
            wD0, wD1 = _word_counts(x0 & D, Wc), _word_counts(x1 & D, Wc)
            f = np.zeros(FEAT, np.float32)
            f[0] = f32((v >> 20) & 63) / f32(depth)
            f[1] = f32(j) * inv_k
            f[2] = (v >> 26) & 1
            f[3] = (v >> 27) & 1
            f[4] = f32(x0.sum()) * inv_lc
            f[5] = f32(x1.sum()) * inv_lc
            f[6] = f32(wD0.sum()) * inv_d
            f[7] = f32(wD1.sum()) * inv_d
            f[8] = f32(_longest_true_run(w0 == 0)) * inv_wc
            f[9] = f32(_longest_true_run(w1 == 0)) * inv_wc
            for o, (a, b) in enumerate(((cw - NEAR, cw), (cw, cw + NEAR), (cw - FAR, cw), (cw, cw + FAR))):
                den = max(wsites(a, b), f32(1))
                f[10 + 2 * o] = rng(w0, a, b) / den
                f[11 + 2 * o] = rng(w1, a, b) / den
            a, b = cw - FAR, cw + FAR
            f[18] = rng(wD0, a, b) / max(rng(wD, a, b), f32(1))
            f[19] = rng(wD1, a, b) / max(rng(wD, a, b), f32(1))
            for o, wv in enumerate((w0, w1)):
                back = 0
                while back < RUN and cw - 1 - back >= 0 and wv[cw - 1 - back] == 0:
                    back += 1
                fwd = 0
                while fwd < RUN and cw + fwd < Wc and wv[cw + fwd] == 0:
                    fwd += 1
                f[20 + o] = f32(back) * f32(1.0 / RUN)
                f[22 + o] = f32(fwd) * f32(1.0 / RUN)
            lo, hi = max(a, 0), min(b, Wc)
            s0, s1 = 64 * lo, min(64 * hi, Lc)
            Rs = R[s0:s1]
            f[24] = f32((Rs & RR[s0:s1]).sum()) / max(rng(wRR, a, b), f32(1))
            f[25] = f32((~Rs & AA[s0:s1]).sum()) / max(rng(wAA, a, b), f32(1))
            f[26] = f32((Rs & RA[s0:s1]).sum()) / max(rng(wRA, a, b), f32(1))
            f[27] = f32(Rs.sum()) / max(wsites(a, b), f32(1))
            f[28] = cid in share0
            f[29] = cid in share1
            f[30] = f32(cw) * inv_wc
            f[31] = wsites(a, b) * f32(1.0 / (64.0 * 2 * FAR))
            X[t, j] = f
    return X


def topk(scores, pool, K, k_learn, states, min_mass):
    """Reference of gh_select_topk: (states, K_out, coverage, kept_pool flags)."""
    T, ks = pool.shape
    st = np.array(states, np.int32, copy=True)
    Ko = np.array(K, np.int32, copy=True)
    cov = np.zeros(T, np.float32)
    kept = np.zeros(T, bool)
    for t in range(T):
        k, Kt = int(k_learn[t]), int(K[t])
        k = min(k, Kt)
        if k <= 0 or k >= Kt:
            continue
        y = scores[t, :Kt].astype(np.float32)
        y = np.where(np.isnan(y), -np.inf, y)
        order = sorted(range(Kt), key=lambda j: (-y[j], j))
        m = np.maximum(np.exp(y[order].astype(np.float64)) - LABEL_EPS, 0.0)
        top, tot = m[:k].sum(), m.sum()
        cov[t] = top / tot if tot > 0 else 0.0
        if not (tot > 0) or top < min_mass * tot:
            kept[t] = True
            continue
        ids = np.sort(pool[t, order[:k]] & 0xFFFFF)
        st[t, :k] = ids
        Ko[t] = k
    return st, Ko, cov, kept


def captured_mass(labels, scores_or_order, k):
    """Share of the true posterior mass (from labels ln(K m + eps)) held by the
    first k candidates of an ordering (an index array) or of a score vector."""
    K = len(labels)
    m = np.maximum((np.exp(np.asarray(labels, np.float64)) - LABEL_EPS) / K, 0.0)
    if scores_or_order.dtype.kind == "f":
        order = np.lexsort((np.arange(K), -scores_or_order))
    else:
        order = scores_or_order
    return m[order[:k]].sum() / max(m.sum(), 1e-300)
