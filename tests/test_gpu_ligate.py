"""GPU parity: gh_ligate / gh_ligate_scan against the ligation oracle
(oracle/ligate_ref.py, restating ligater_algorithm.cpp:90-216, 339-444):
emission order, per-chunk swaps and the ligated genotypes are identical."""
import numpy as np
import pytest

import ligate_ref as lr
from conc_cases import lig_case, VECTOR_END

pytestmark = pytest.mark.gpu


def _roughen(chunks, seed):
    """Unphased, haploid and multi-allelic entries the scan must skip."""
    rng = np.random.default_rng(seed)
    for c in chunks:
        gt = c["gt"]
        u = rng.random(gt.shape[:2]) < 0.03
        gt[u, 1] &= ~1                      # unphased
        h = rng.random(gt.shape[:2]) < 0.01
        gt[h, 1] = VECTOR_END               # haploid
        c["biallelic"][rng.random(len(c["pos"])) < 0.05] = 0


def _gather(chunks, order):
    return np.stack([chunks[c]["gt"][r] for c, r in order])


@pytest.mark.parametrize("n_chunks,N,L,ov,seed,brk,rough", [
    (2, 50, 300, 80, 1, None, False),
    (5, 37, 200, 60, 2, None, True),
    (4, 300, 150, 40, 3, 2, True),
    (3, 1000, 120, 50, 4, None, True),
])
def test_ligate_matches_oracle(n_chunks, N, L, ov, seed, brk, rough):
    from glimpse_amd import ligate as gl
    chunks, truth, flips = lig_case(n_chunks, N, L, ov, seed=seed, contig_break=brk)
    if rough:
        _roughen(chunks, seed)
    ref_chunks = [dict(c, gt=c["gt"].copy()) for c in chunks]
    want, want_order, want_swaps = lr.ligate(ref_chunks, N)
    order, swap, stats = gl.ligate(chunks, N)
    assert [tuple(x) for x in order.tolist()] == [tuple(x) for x in want_order]
    for f in range(n_chunks):
        assert np.array_equal(swap[f].astype(bool), want_swaps[f]), "chunk %d" % f
    assert np.array_equal(_gather(chunks, order), want)
    assert len(stats) == n_chunks - 1
    assert sum(s["n_swapped"] for s in stats) > 0


def test_ligate_scan_matches_update_distances():
    from glimpse_amd import ligate as gl
    chunks, truth, flips = lig_case(2, 200, 100, 100, seed=7)
    _roughen(chunks, 7)
    a, b = chunks[0]["gt"], chunks[1]["gt"]
    swap_a = (np.random.default_rng(7).random(200) < 0.5).astype(np.uint8)
    nmatch, nmism, swap_b, q = gl.scan(a, b, swap_a)
    wm, wx = [0] * 200, [0] * 200
    for r in range(a.shape[0]):
        lr.update_distances(a[r], b[r], swap_a.astype(bool), wm, wx)
    assert nmatch.tolist() == wm and nmism.tolist() == wx
    assert np.array_equal(swap_b.astype(bool), np.array(wm) < np.array(wx))
    wq = np.array([lr.phase_quality(m, x) for m, x in zip(wm, wx)], np.float32)
    assert np.allclose(q, wq, atol=1e-4)
