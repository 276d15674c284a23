"""gh_impute_batch_noalloc (the forward-backward batch with its device buffers
in a file-scope scratch array) against gh_impute_batch on the same inputs:
two calls in a row, the second smaller, through the C-ABI; identical kernels,
so identical posteriors.  A batch larger than the scratch array is refused
before anything runs."""
import ctypes

import numpy as np
import pytest

import orc
from glimpse_amd import hip, synth

pytestmark = pytest.mark.gpu


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _csr(sets):
    ptr = np.zeros(len(sets) + 1, np.int32)
    ptr[1:] = np.cumsum([len(s) for s in sets])
    return ptr, np.concatenate([np.asarray(s, np.int32) for s in sets])


def _noalloc(panel, sets, HL, flat):
    ptr, ids = _csr(sets)
    n = len(sets)
    HL = np.ascontiguousarray(HL, np.float32)
    flat = np.ascontiguousarray(flat, np.uint8)
    HP = np.full((n, 2 * panel.L), np.nan, np.float32)
    rc = hip.lib().gh_impute_batch_noalloc(panel.h, n, _p(ptr), _p(ids), _p(HL), _p(flat), ctypes.c_float(1e-12),
                                           100000, _p(HP))
    return rc, HP


@pytest.fixture(scope="module")
def case():
    P = synth.make_panel(2000, 3000, 4, length_mb=1.0, rare_frac=0.3, flip_frac=0.05, seed=77)
    gl, flat = synth.make_reads(P["tgt_haps"], 1.0, seed=78)
    return P, gl, flat, hip.Panel(P, with_pbwt=False)


def _inputs(P, gl, flat, Ks, rng):
    sets = [np.sort(rng.choice(P["N"], K, replace=False)) for K in Ks]
    T = flat.shape[0]
    HL = np.stack([orc.hap_likelihoods(gl[i % T], flat[i % T]) for i in range(len(Ks))])
    fl = np.stack([flat[i % T] for i in range(len(Ks))])
    return sets, HL, fl


def test_noalloc_matches_impute_batch_twice(case):
    P, gl, flat, GP = case
    rng = np.random.default_rng(9)
    for Ks in ([2000, 700, 64, 1, 300, 1500], [100, 37]):
        sets, HL, fl = _inputs(P, gl, flat, Ks, rng)
        want = GP.impute_batch(sets, HL, fl)
        rc, got = _noalloc(GP, sets, HL, fl)
        assert rc == 0, hip.lib().gh_last_error().decode()
        assert np.array_equal(got, want), "K=%s: max |diff| %g" % (Ks, np.nanmax(np.abs(got - want)))


def test_noalloc_refuses_batches_beyond_the_scratch(case):
    P, gl, flat, GP = case
    rng = np.random.default_rng(10)
    sets, HL, fl = _inputs(P, gl, flat, [2000] * 64, rng)
    rc, _ = _noalloc(GP, sets, HL, fl)
    assert rc == -4 and "scratch" in hip.lib().gh_last_error().decode()
