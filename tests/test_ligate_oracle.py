"""Pins of the ligation oracle (oracle/ligate_ref.py) on the CPU: chunks cut
from one phased truth, with the phase of random samples flipped per chunk,
must be aligned back to the truth (ligater_algorithm.cpp:108-127, 196-213,
90-106), every site emitted once in order (:339-444); the phase quality
against its closed form."""
import math

import numpy as np
import pytest

import ligate_ref as lr
from conc_cases import lig_case


def _alleles(gt):
    return (np.asarray(gt, np.int64) >> 1) - 1


@pytest.mark.parametrize("n_chunks,N,L,ov,seed", [(2, 20, 120, 40, 1), (4, 16, 100, 30, 2), (3, 40, 80, 38, 3)])
def test_oracle_undoes_chunk_flips(n_chunks, N, L, ov, seed):
    chunks, truth, flips = lig_case(n_chunks, N, L, ov, seed=seed)
    out, order, swaps = lr.ligate(chunks, N)
    # every truth site once, in order
    glob = [chunks[c]["lo"] + r for c, r in order]
    assert glob == list(range(truth.shape[0]))
    # the chunk swaps are the flips (each overlap holds hets for every sample here)
    for f in range(n_chunks):
        assert np.array_equal(swaps[f], flips[f]), "chunk %d" % f
    a_out, a_tru = _alleles(out), _alleles(truth)
    ok = (a_tru[..., 0] >= 0) & (a_tru[..., 1] >= 0)
    assert np.array_equal(a_out[ok], a_tru[ok])
    # the second allele stays phased
    assert np.all(out[ok][:, 1] & 1)


def test_first_half_of_overlap_from_the_earlier_chunk():
    chunks, truth, flips = lig_case(2, 8, 100, 40, seed=4)
    out, order, swaps = lr.ligate(chunks, 8)
    src = [c for c, r in order]
    # 60 records of chunk 0 alone, 20 of its overlap records, then 20 + 60 of chunk 1
    assert src == [0] * 80 + [1] * 80


def test_no_overlap_carries_the_swap():
    chunks, truth, flips = lig_case(3, 10, 60, 20, seed=5, contig_break=2)
    out, order, swaps = lr.ligate(chunks, 10)
    assert np.array_equal(swaps[1], flips[1])
    assert np.array_equal(swaps[2], swaps[1])           # another contig: chunk 2 keeps chunk 1's swaps
    assert len(order) == 60 + 40 + 60                   # chunk 2's first 20 sites are emitted again


def test_three_overlapping_chunks_refused():
    chunks, truth, flips = lig_case(3, 6, 80, 60, seed=6)   # chunk 2 starts inside chunk 0
    with pytest.raises(ValueError):
        lr.ligate(chunks, 6)


def test_library_refuses_three_overlapping_chunks():
    """The library's host planning refuses the case too, before any device work."""
    from glimpse_amd import hip
    from glimpse_amd import ligate as gl
    chunks, truth, flips = lig_case(3, 6, 80, 60, seed=6)
    with pytest.raises(hip.GhError, match="three chunks overlap"):
        gl.ligate(chunks, 6)


def test_phase_quality_closed_form():
    assert lr.phase_quality(5, 0) == np.float32(99)
    assert lr.phase_quality(0, 7) == np.float32(99)
    f0 = 0.75
    want = 99 * (0.7 + f0 * math.log(f0) + (1 - f0) * math.log(1 - f0)) / 0.7
    assert abs(float(lr.phase_quality(3, 1)) - want) < 1e-4
    # maximal uncertainty: 99 (0.7 - ln 2) / 0.7
    assert abs(float(lr.phase_quality(4, 4)) - 99 * (0.7 - math.log(2)) / 0.7) < 1e-4


def test_phase_update_rules():
    V = lr.VECTOR_END
    gt = np.array([[2, 5], [4, 3], [0, 5], [4, V], [2, 4], [4, 5]], np.int64)   # 0|1, 1|0, .|1, haploid, 0/1, 1|1
    lr.phase_update(gt, np.ones(6, bool))
    assert gt.tolist() == [[5, 3], [3, 5], [0, 5], [4, V], [2, 4], [5, 5]]
