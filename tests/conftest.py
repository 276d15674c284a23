import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libglimpse_hip.so)")


def hvar_bytes_to_words(hvar_bytes, K):
    """Oracle Hvar rows (MSB-first bytes, ceil(K/8) per row) -> uint64 words
    (bit j of word w = state 64w+j), padding bits cleared."""
    rows = hvar_bytes.shape[0]
    bits = np.unpackbits(hvar_bytes, axis=1, bitorder="big")[:, :K]
    W = (K + 63) // 64
    pad = np.zeros((rows, W * 64), np.uint8)
    pad[:, :K] = bits
    return np.packbits(pad.reshape(rows, W, 64), axis=2, bitorder="little").view(np.uint64).reshape(rows, W)


def mask_words(words, K):
    W = words.shape[1]
    out = words.copy()
    rem = K - 64 * (W - 1)
    if rem < 64:
        out[:, W - 1] &= np.uint64((1 << rem) - 1)
    return out


@pytest.fixture(scope="session")
def small_case():
    """N=2000 haplotypes x 3000 sites with injected rare variants, 4 targets at 1x."""
    from glimpse_amd import synth
    P = synth.make_panel(2000, 3000, 4, length_mb=1.0, rare_frac=0.3, flip_frac=0.05, seed=11)
    gl, flat = synth.make_reads(P["tgt_haps"], 1.0, seed=5)
    return P, gl, flat


def dosage(gp0, gp1):
    """DS = gp1 + 2 gp2 with gp2 = 1 - gp0 - gp1."""
    return gp1 + 2.0 * np.clip(1.0 - gp0 - gp1, 0.0, 1.0)


def dosage_r2(gp0, gp1, truth):
    """Squared Pearson correlation of DS against the true genotype over every
    target x site pair (concordance/src/containers/call_set_reading.cpp:688-693,
    common/src/utils/basic_stats.h:61-125)."""
    c = np.corrcoef(dosage(gp0, gp1).ravel().astype(np.float64), np.asarray(truth, np.float64).ravel())[0, 1]
    return float(c * c)


def nrc(g, truth):
    """Non-reference concordance 1 - NRD of best-guess genotypes: NRD = mismatches
    / (mismatches + matches at RA and AA truths)
    (concordance/src/containers/call_set_writing.cpp:46-55)."""
    g = np.asarray(g, np.int32).ravel()
    t = np.asarray(truth, np.int32).ravel()
    mism = int(np.sum(g != t))
    match_nonref = int(np.sum((g == t) & (t > 0)))
    return 1.0 - mism / max(1, mism + match_nonref)


def haploid_gls(P, targets, cov=2.0, seed=3):
    """Diploid-shaped GL arrays where the listed targets are haploid: their
    GLs are (PL of the hom-ref, PL of the hom-alt) of a homozygous read set,
    rescaled to a minimum of 0."""
    from glimpse_amd import synth
    haps = P["tgt_haps"].copy()
    for t in targets:
        haps[2 * t + 1] = haps[2 * t]
    gl, flat = synth.make_reads(haps, cov, seed=seed)
    for t in targets:
        g = np.stack([gl[t, :, 0], gl[t, :, 2]], -1).astype(np.int32)
        g -= g.min(axis=1, keepdims=True)
        gl[t, :, 0], gl[t, :, 1], gl[t, :, 2] = g[:, 0], g[:, 1], 0
        flat[t] |= (g[:, 0] == g[:, 1]).astype(np.uint8)
    return gl, flat
