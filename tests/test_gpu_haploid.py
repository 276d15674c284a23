"""GPU parity for haploid targets (ploidy 1: one forward-backward pass per
iteration on the marginal haplotype likelihoods, one PBWT walker, no
phasing, haploid GP storage and output; caller_algorithm.cpp:53-84,
genotype.cpp:56-90, 127-166, genotype_writer.cpp:138-190) against the oracle
in counter RNG mode, in runs that mix haploid and diploid targets."""
import numpy as np
import pytest

import orc
from conftest import haploid_gls
from glimpse_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case():
    from glimpse_amd import hip
    P = synth.make_panel(1200, 2500, 6, length_mb=0.5, rare_frac=0.3, seed=41)
    P["lq"] = (np.arange(P["L"]) % 13 == 4).astype(np.uint8)
    return P, orc.Panel(P), hip.Panel(P)


def _compare(got, ref, ploidy):
    assert np.array_equal(got["nstates"], ref["nstates"]), "conditioning-set sizes differ"
    d = np.maximum(np.abs(got["gp0"] - ref["gp0"]), np.abs(got["gp1"] - ref["gp1"]))
    assert np.mean(d < 1e-4) > 0.995, "GP differs at %.4f of sites" % np.mean(d >= 1e-4)
    hap = np.asarray(ploidy) == 1
    assert np.mean(got["h0"][hap] == ref["h0"][hap]) > 0.995
    assert np.all(got["h1"][hap] == 0) and np.all(got["hds"][hap] == 0)


@pytest.mark.parametrize("ploidy,kpbwt", [([1, 2, 2, 1, 2, 1], 300), ([1, 1, 1, 1, 1, 1], 500),
                                          ([2, 1, 2, 2, 2, 2], 2000)])
def test_haploid_run_matches_oracle(case, ploidy, kpbwt):
    from glimpse_amd import hip
    P, OP, GP = case
    hap_t = [t for t, p in enumerate(ploidy) if p == 1]
    gl, flat = haploid_gls(P, hap_t, cov=1.0, seed=kpbwt)
    prm_o = orc.default_params(kpbwt=kpbwt, burnin=1, main_iter=2, rng_mode=orc.RNG_COUNTER)
    ref = OP.phase_run(gl, flat, prm_o, ploidy=ploidy)
    prm = hip.default_params(kpbwt=kpbwt, burnin=1, main_iter=2)
    r = hip.Run(GP, gl, flat, prm, ploidy=ploidy)
    r.execute()
    got = r.fetch()
    _compare(got, ref, ploidy)
    st = got["stats"] if "stats" in got else r.stats()
    n_passes = sum(ploidy)
    assert st["fb_poly_sites"] > 0
    r.close()


def test_haploid_output_fields_bit_exact():
    from glimpse_amd import hip
    rng = np.random.default_rng(5)
    T, L = 9, 2000
    pl = np.array([1, 2, 1, 1, 2, 2, 1, 2, 1], np.int32)
    a = rng.dirichlet([0.3, 0.3, 0.3], (T, L)).astype(np.float32)
    b = rng.dirichlet([0.3, 0.3], (T, L)).astype(np.float32)
    gp0 = np.where(pl[:, None] == 1, b[..., 0], a[..., 0]).astype(np.float32)
    gp1 = np.where(pl[:, None] == 1, b[..., 1], a[..., 1]).astype(np.float32)
    gp0[:, :50], gp1[:, :50] = 1.0, 0.0
    hds = rng.integers(0, 2, (T, L)).astype(np.uint8)
    cref = rng.integers(1, 500, L).astype(np.uint32)
    calt = rng.integers(1, 500, L).astype(np.uint32)
    for ploidy in (pl, np.ones(T, np.int32)):
        g0 = np.where(ploidy[:, None] == 1, gp0, a[..., 0]).astype(np.float32)
        g1 = np.where(ploidy[:, None] == 1, gp1, a[..., 1]).astype(np.float32)
        want = orc.output_fields(g0, g1, hds, cref, calt, ploidy=ploidy)
        got = hip.output_fields(g0, g1, hds, cref, calt, ploidy=ploidy)
        for k in ("ds", "gt", "info", "af", "raf"):
            assert np.array_equal(got[k], want[k]), k
        assert np.array_equal(got["gp"], want["gp"], equal_nan=True)
