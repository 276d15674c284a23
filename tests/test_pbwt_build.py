"""Sparse-PBWT build (split_reference side): the product encoder must produce
the oracle's pack3 stream and A_small_idx byte for byte, and the oracle's
stream must decode to Durbin's prefix arrays (brute force on small panels)."""
import numpy as np
import pytest

import orc
from glimpse_amd import hip, synth


def _panel(N, L, rare_frac, seed, lq_every=0):
    P = synth.make_panel(N, L, 2, length_mb=L / 5000.0, rare_frac=rare_frac, flip_frac=0.05, seed=seed,
                         keep_dense=True)
    if lq_every:
        P["lq"] = (np.arange(P["L"]) % lq_every == 1).astype(np.uint8)
    return P


@pytest.mark.parametrize("N,L,rare,lq", [(300, 800, 0.0, 0), (2000, 1500, 0.4, 0), (2500, 1200, 0.6, 5)])
def test_product_pbwt_stream_matches_oracle(N, L, rare, lq):
    P = _panel(N, L, rare, seed=N + L, lq_every=lq)
    y, ap, ai = hip.build_sparse_pbwt(P)
    OP = orc.Panel(P)
    assert np.array_equal(y, OP.ypacked())
    optr, oidx = OP.asmall()
    assert np.array_equal(ap, optr) and np.array_equal(ai, oidx)


def _decode_common_only(y, N, L):
    """Decode a stream of an all-common, all-HQ panel into the prefix arrays."""
    A = np.arange(N)
    pos = 0
    arrays = []
    cols = []
    for _ in range(L):
        m = 0
        syms = []
        while m < N:
            z = int(y[pos]); pos += 1
            c = z & 0x7F
            n = c if c < 64 else ((c - 64) << 6 if c < 96 else (c - 96) << 11)
            syms += [z >> 7] * n
            m += n
        syms = np.array(syms)
        cols.append(syms.copy())
        A = np.concatenate([A[syms == 0], A[syms == 1]])
        arrays.append(A.copy())
    assert pos == len(y)
    return arrays, cols


def test_oracle_stream_is_durbin_pbwt():
    """Brute force (Durbin 2014): after site k the prefix array sorts haplotypes
    by their reversed prefix (ties broken by index); the stored column is the
    site's alleles in the previous order."""
    N, L = 64, 120
    P = _panel(N, L, 0.0, seed=3)
    assert P["common"].all()
    OP = orc.Panel(P)
    arrays, cols = _decode_common_only(OP.ypacked(), N, P["L"])
    X = P["dense"]  # [N][L]
    prev = np.arange(N)
    for k in range(P["L"]):
        keys = [tuple(X[h, k::-1]) + (h,) for h in range(N)]
        want = np.array(sorted(range(N), key=lambda h: keys[h]))
        assert np.array_equal(arrays[k], want), "prefix array differs at site %d" % k
        assert np.array_equal(cols[k], X[prev, k]), "column order differs at site %d" % k
        prev = arrays[k]


def _pack3(sym, n):
    y, codes = sym << 7, []
    while n >= 63488:
        codes.append(y | 0x7F); n -= 63488
    if n >= 2048:
        codes.append(y | 0x60 | (n >> 11)); n &= 0x7FF
    if n >= 64:
        codes.append(y | 0x40 | (n >> 6)); n &= 0x3F
    if n:
        codes.append(y | n)
    return codes


def test_pack3_code_lengths():
    """Closed-form pack3 codes (ref_haplotype_set.h:45-85) for long runs:
    69900 zeros then 100 ones -> 0x7f, 0x63, 0x44, 0x0c, 0xc1, 0xa4."""
    N, ones = 70000, 100
    P = dict(L=1, N=N, cm=np.zeros(1), cref=np.array([N - ones], np.uint32), calt=np.array([ones], np.uint32),
             lq=np.zeros(1, np.uint8))
    W = (N + 63) // 64
    words = np.zeros((1, W), np.uint64)
    for h in range(N - ones, N):
        words[0, h >> 6] |= np.uint64(1) << np.uint64(h & 63)
    P.update(common_words=words, rare_ptr=np.zeros(2, np.int32), rare_idx=np.zeros(0, np.int32))
    y, _, _ = hip.build_sparse_pbwt(P)
    assert list(y) == _pack3(0, N - ones) + _pack3(1, ones)
    assert list(y) == [0x7F, 0x63, 0x44, 0x0C, 0xC1, 0xA4]
