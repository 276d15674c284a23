"""Seeded synthetic inputs for the concordance and ligation tests (truth and
estimate records as the concordance tool reads them from VCF/BCF, phased
chunks as the ligate tool reads them)."""
import numpy as np

MISSING = -2147483648
VECTOR_END = -2147483647


def conc_case(S, N, ploidy=2, seed=1, miss=0.03, indel_frac=0.2, alt_af=False):
    """Truth PL / GT / DP and estimate GP / DS / GT for S biallelic sites and N
    samples.  Estimates are noisy posteriors around the truth; a share of every
    field is missing; some sites lack AF or are of another type."""
    rng = np.random.default_rng(seed)
    P = ploidy
    af = rng.beta(0.3, 2.0, S).astype(np.float32)
    flip_sites = rng.random(S) < 0.2
    af[flip_sites] = 1.0 - af[flip_sites]          # some alt-major sites (flipped by the tool)
    af[rng.random(S) < 0.02] = np.nan               # no AF: site skipped
    af[:3] = np.array([0.5, 0.0, 1.0], np.float32)[: min(3, S)]
    type_ = np.where(rng.random(S) < indel_frac, 2, 1).astype(np.uint8)
    type_[rng.random(S) < 0.02] = 0                 # another variant type
    p = np.nan_to_num(af, nan=0.3).clip(0.001, 0.999)
    g = (rng.random((S, N, P)) < p[:, None, None]).sum(axis=2).astype(np.int8)   # true genotype
    # truth PL: 0 at the true genotype, 3..60 elsewhere; some low-confidence ones
    pl = rng.integers(3, 61, (S, N, P + 1)).astype(np.int32)
    pl[np.arange(S)[:, None], np.arange(N)[None, :], g] = 0
    weak = rng.random((S, N)) < 0.05
    pl[weak] = rng.integers(0, 3, (int(weak.sum()), P + 1))
    big = rng.random((S, N)) < 0.01
    pl[big, 0] = 300                                # above 255: capped
    m = rng.random((S, N)) < miss
    pl[m, rng.integers(0, P + 1, int(m.sum()))] = MISSING
    m = rng.random((S, N)) < miss / 3
    pl[m, P] = VECTOR_END
    tgt = g.copy()
    tgt[rng.random((S, N)) < miss] = -1
    dp = rng.integers(0, 25, (S, N)).astype(np.int32)
    dp[rng.random((S, N)) < miss] = MISSING
    # estimates: posterior mass around the truth, with errors
    wrong = rng.random((S, N)) < 0.08
    e = np.where(wrong, rng.integers(0, P + 1, (S, N)), g).astype(np.int64)
    gp = rng.random((S, N, P + 1)).astype(np.float32) ** 3
    gp[np.arange(S)[:, None], np.arange(N)[None, :], e] += rng.uniform(0.5, 4.0, (S, N)).astype(np.float32)
    gp /= gp.sum(axis=2, keepdims=True)
    r = rng.random((S, N)) < 0.5                     # half of them rounded as the writer does
    gp[r] = np.floor(gp[r] * 1000.0) / 1000.0
    gp = gp.astype(np.float32)
    ties = rng.random((S, N)) < 0.01
    gp[ties] = np.array([0.5, 0.5, 0.0][: P + 1] if P == 2 else [0.5, 0.5], np.float32)
    gp[rng.random((S, N)) < miss / 2, 0] = np.nan
    ds = (gp[:, :, 1] + (2.0 * gp[:, :, 2] if P == 2 else 0.0)).astype(np.float32)
    egt = e.astype(np.int8)
    egt[rng.random((S, N)) < miss / 2] = -1
    return {"af": af, "type": type_, "truth_pl": pl, "truth_gt": tgt, "truth_dp": dp, "esti_gp": gp,
            "esti_ds": ds, "esti_gt": egt, "ploidy": P}


def phased(a0, a1):
    """htslib GT values of phased diploid genotypes as a writer stores them
    ("a0|a1": the first allele without the phase bit, the second with it)."""
    return np.stack([(a0 + 1) << 1, ((a1 + 1) << 1) | 1], axis=-1).astype(np.int32)


def lig_case(n_chunks, N, L_chunk, overlap, seed=1, flip_p=0.4, miss=0.02, contig_break=None):
    """A phased truth cut into overlapping chunks; each chunk after the first
    has the phase of a random set of samples flipped (relative to the truth).
    Returns (chunks, truth rows [L][N][2], flips [n_chunks][N])."""
    rng = np.random.default_rng(seed)
    step = L_chunk - overlap
    L = step * (n_chunks - 1) + L_chunk
    pos = np.cumsum(rng.integers(1, 50, L)).astype(np.int32)
    key = rng.integers(0, 2**62, L).astype(np.uint64)
    h = (rng.random((L, N, 2)) < rng.uniform(0.05, 0.5, L)[:, None, None]).astype(np.int32)
    h[rng.random((L, N)) < miss, 0] = -1   # a missing allele
    truth = phased(h[..., 0], h[..., 1])
    chunks, flips = [], np.zeros((n_chunks, N), bool)
    for f in range(n_chunks):
        lo, hi = f * step, f * step + L_chunk
        contig = np.zeros(hi - lo, np.int32)
        if contig_break is not None and f >= contig_break:
            contig[:] = 1
        fl = rng.random(N) < flip_p if f > 0 else np.zeros(N, bool)
        flips[f] = fl
        hh = h[lo:hi].copy()
        hh[:, fl, 0], hh[:, fl, 1] = h[lo:hi][:, fl, 1], h[lo:hi][:, fl, 0]
        chunks.append({"contig": contig, "pos": pos[lo:hi].copy(), "key": key[lo:hi].copy(),
                       "biallelic": np.ones(hi - lo, np.uint8), "gt": phased(hh[..., 0], hh[..., 1]), "lo": lo})
    return chunks, truth, flips
