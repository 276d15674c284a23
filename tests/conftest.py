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
