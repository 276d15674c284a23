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


def features(P, H, gl, flat, pool, K, depth):
    """X [T][kstride][32] for the pools (rows >= K[t] stay zero)."""
    T, ks = pool.shape
    L = P["L"]
    common = P["common"].astype(bool)
    c2t = np.nonzero(common)[0]
    Lc = len(c2t)
    Wc = (Lc + 63) // 64
    A = _common_alleles(P)
    major = P["major"].astype(np.uint8)
    rp, ri = P["rare_ptr"], P["rare_idx"]
    X = np.zeros((T, ks, FEAT), np.float32)
    word_of = np.arange(Lc) // 64
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
        inv_rr = f32(1.0) / f32(max(RR.sum(), 1))
        inv_aa = f32(1.0) / f32(max(AA.sum(), 1))
        inv_ra = f32(1.0) / f32(max(RA.sum(), 1))
        inv_lc = f32(1.0) / f32(max(Lc, 1))
        inv_wc = f32(1.0) / f32(max(Wc, 1))
        # panel haplotypes carrying a rare allele that H0 / H1 carries
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
            R = A[:, cid]
            x0, x1 = R ^ h0, R ^ h1
            f = np.zeros(FEAT, np.float32)
            f[0] = f32((v >> 20) & 63) / f32(depth)
            f[1] = f32(j) * inv_k
            f[2] = (v >> 26) & 1
            f[3] = (v >> 27) & 1
            f[4] = f32((x0 & D).sum()) * inv_d
            f[5] = f32((x1 & D).sum()) * inv_d
            f[6] = f32(x0.sum()) * inv_lc
            f[7] = f32(x1.sum()) * inv_lc
            clean0 = np.bincount(word_of, weights=x0, minlength=Wc) == 0
            clean1 = np.bincount(word_of, weights=x1, minlength=Wc) == 0
            f[8] = f32(_longest_true_run(clean0)) * inv_wc
            f[9] = f32(_longest_true_run(clean1)) * inv_wc
            f[10] = f32((R & RR).sum()) * inv_rr
            f[11] = f32((~R & AA).sum()) * inv_aa
            f[12] = f32((R & RA).sum()) * inv_ra
            f[13] = f32(R.sum()) * inv_lc
            f[14] = cid in share0
            f[15] = cid in share1
            for w in range(WIN):
                ws, we = (w * Wc) // WIN, ((w + 1) * Wc) // WIN
                s0, s1 = 64 * ws, min(64 * we, Lc)
                nsite = f32(max(s1 - 64 * ws, 1))
                f[16 + w] = f32(x0[s0:s1].sum()) / nsite
                f[24 + w] = f32(x1[s0:s1].sum()) / nsite
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
