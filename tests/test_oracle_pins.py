"""Pins of the oracle's sampling, INIT selection, GP storage and random draws
(no GPU), each against what the mathematics or a published algorithm fixes:

* INIT sampling (haplotype_set.cpp:775-833): std::sample (reference mode) and
  Knuth's Algorithm S with counter draws (counter mode) both draw k of n
  without replacement with every k-subset equally likely, in population order;
* INIT selection: an all-flat target gets a uniform Kinit-subset; a target
  with certain evidence for the rare allele at one site keeps every carrier of
  that site (size_max = floor(0.8 Kinit / #sites) >= carriers);
* GP storage and finalisation (genotype.cpp:168-244) against a dense closed
  form: a site contributes the implicit (1, 0) until the first iteration whose
  P(RR) < 0.99999, then its normalised GP, all divided by the iteration count;
* getFloat: reference mode is libstdc++'s generate_canonical<float, 24> of one
  mt19937 output (x / 2^32 in float; numpy's legacy MT19937 gives the same
  stream), counter mode Philox draws are uniform (chi-square) and the haplotype
  sampler H = u > P(allele 0) picks allele 1 with probability P(allele 1)."""
import math

import numpy as np
import pytest

import orc
from glimpse_amd import synth


def _chi2_sf(x, dof):
    """Upper tail of the chi-square distribution (Wilson-Hilferty)."""
    z = ((x / dof) ** (1.0 / 3.0) - (1.0 - 2.0 / (9.0 * dof))) / math.sqrt(2.0 / (9.0 * dof))
    return 0.5 * math.erfc(z / math.sqrt(2.0))


@pytest.mark.parametrize("mode", [orc.RNG_REFERENCE, orc.RNG_COUNTER])
def test_sampling_without_replacement_is_uniform(mode):
    n, k, reps = 20, 5, 4000
    counts = np.zeros(n)
    for r in range(reps):
        s = orc.sample(n, k, mode, seed=1000 + r, tgt=r % 7)
        assert len(s) == k and len(set(s.tolist())) == k
        assert np.all(np.diff(s) > 0), "sample not in population order"
        counts[s] += 1
    exp = reps * k / n
    chi2 = float(np.sum((counts - exp) ** 2 / exp))
    assert _chi2_sf(chi2, n - 1) > 1e-3, "inclusion frequencies chi2 %.1f" % chi2
    # every 2-subset of 6 equally likely
    sub = {}
    for r in range(6000):
        s = tuple(orc.sample(6, 2, mode, seed=77 + r, tgt=3).tolist())
        sub[s] = sub.get(s, 0) + 1
    assert len(sub) == 15
    exp = 6000 / 15
    chi2 = sum((c - exp) ** 2 / exp for c in sub.values())
    assert _chi2_sf(chi2, 14) > 1e-3, "subset frequencies chi2 %.1f" % chi2


def test_sampling_closed_cases():
    for mode in (orc.RNG_REFERENCE, orc.RNG_COUNTER):
        assert np.array_equal(orc.sample(7, 7, mode, 5), np.arange(7))
        assert np.array_equal(orc.sample(7, 20, mode, 5), np.arange(7))
        assert len(orc.sample(7, 0, mode, 5)) == 0


@pytest.fixture(scope="module")
def init_case():
    P = synth.make_panel(3000, 1500, 2, length_mb=0.5, rare_frac=0.4, flip_frac=0.05, seed=91)
    return P, orc.Panel(P)


@pytest.mark.parametrize("mode", [orc.RNG_REFERENCE, orc.RNG_COUNTER])
def test_init_selection_flat_target_is_uniform(init_case, mode):
    P, OP = init_case
    N, L, kinit, reps = P["N"], P["L"], 100, 600
    gl = np.zeros((L, 3), np.uint8)
    flat = np.ones(L, np.uint8)
    counts = np.zeros(N)
    for r in range(reps):
        s = OP.init_states(gl, flat, kinit, mode, seed=500 + r, tgt=r % 5)
        assert len(s) == kinit and np.all(np.diff(s) > 0)
        counts[s] += 1
    exp = reps * kinit / N
    chi2 = float(np.sum((counts - exp) ** 2 / exp))
    assert _chi2_sf(chi2, N - 1) > 1e-3, "chi2 %.1f" % chi2
    assert len(OP.init_states(gl, flat, 5000, mode, seed=1)) == N   # Kinit clipped to N


@pytest.mark.parametrize("mode", [orc.RNG_REFERENCE, orc.RNG_COUNTER])
def test_init_selection_keeps_carriers_of_a_certain_rare_allele(init_case, mode):
    P, OP = init_case
    L = P["L"]
    rare = np.nonzero(~P["common"].astype(bool))[0]
    ncar = np.diff(P["rare_ptr"])[rare]
    assert np.any(ncar >= 2)
    v = int(rare[np.argmax(ncar >= 2)])
    carriers = P["rare_idx"][P["rare_ptr"][v]:P["rare_ptr"][v + 1]]
    gl = np.zeros((L, 3), np.uint8)
    flat = np.ones(L, np.uint8)
    flat[v] = 0
    # certain heterozygote: the only site with evidence for a rare allele
    gl[v] = (255, 0, 255)
    for r in range(20):
        s = set(OP.init_states(gl, flat, 50, mode, seed=900 + r).tolist())
        assert len(s) == 50
        assert set(carriers.tolist()) <= s, "carriers of the rare site missing"


def _gp_closed_form(HP0, HP1):
    n_iter, L2 = HP0.shape
    L = L2 // 2
    a0, a1 = HP0[:, 0::2].astype(np.float64), HP0[:, 1::2].astype(np.float64)
    b0, b1 = HP1[:, 0::2].astype(np.float64), HP1[:, 1::2].astype(np.float64)
    p0, p1, p2 = a0 * b0, a0 * b1 + a1 * b0, a1 * b1
    n0, n1 = p0 / (p0 + p1 + p2), p1 / (p0 + p1 + p2)
    low = n0 < 0.99999
    stored = low.any(axis=0)
    first = np.where(stored, np.argmax(low, axis=0), n_iter)
    it = np.arange(n_iter)[:, None]
    after = it >= first[None, :]
    gp0 = (first + np.sum(np.where(after, n0, 0.0), axis=0)) / n_iter
    gp1 = np.sum(np.where(after, n1, 0.0), axis=0) / n_iter
    gp0 = np.where(stored, gp0, 1.0)
    gp1 = np.where(stored, gp1, 0.0)
    hds = np.where(stored, a1[-1] < b1[-1], False)
    return gp0, gp1, hds, stored


def test_gp_storage_closed_form():
    rng = np.random.default_rng(4)
    n_iter, L = 15, 4000
    q0 = rng.uniform(0.0, 1.0, (n_iter, L))
    q1 = rng.uniform(0.0, 1.0, (n_iter, L))
    # a third of the sites stay certain RR for a random number of iterations
    certain = rng.random(L) < 0.33
    until = rng.integers(0, n_iter + 1, L)
    it = np.arange(n_iter)[:, None]
    hold = certain[None, :] & (it < until[None, :])
    q0 = np.where(hold, 1e-7, q0)
    q1 = np.where(hold, 1e-7, q1)
    HP0 = np.empty((n_iter, 2 * L), np.float32)
    HP1 = np.empty((n_iter, 2 * L), np.float32)
    HP0[:, 0::2], HP0[:, 1::2] = 1 - q0, q0
    HP1[:, 0::2], HP1[:, 1::2] = 1 - q1, q1
    got = orc.gp_run(HP0, HP1)
    gp0, gp1, hds, stored = _gp_closed_form(HP0, HP1)
    assert (~stored).sum() > 40 and stored.sum() > 1000
    assert np.abs(got["gp0"] - gp0).max() < 2e-6 and np.abs(got["gp1"] - gp1).max() < 2e-6
    assert np.array_equal(got["hds"].astype(bool), hds)
    g2 = 1.0 - got["gp0"] - got["gp1"]
    gt = np.where((got["gp0"] > got["gp1"]) & (got["gp0"] > g2), 0,
                  np.where((got["gp1"] > got["gp0"]) & (got["gp1"] > g2), 1,
                           np.where((g2 > got["gp0"]) & (g2 > got["gp1"]), 2, 0)))
    assert np.array_equal(got["h0"].astype(int) + got["h1"], gt)


def test_reference_draws_are_generate_canonical_of_mt19937():
    for seed in (5489, 15052011):
        raw = np.random.RandomState(seed)._bit_generator.random_raw(5000).astype(np.uint64)
        want = (raw.astype(np.float32) / np.float32(2.0 ** 32)).astype(np.float32)
        want = np.where(want >= 1.0, np.nextafter(np.float32(1), np.float32(0)), want)
        got = orc.draws(seed, 0, 0, 0, 5000, orc.RNG_REFERENCE)
        assert np.array_equal(got, want)


def test_counter_draws_are_uniform_and_sample_alleles():
    u = np.concatenate([orc.draws(11, it, 0, t, 50000, orc.RNG_COUNTER) for it in range(2) for t in range(2)])
    assert u.min() >= 0.0 and u.max() < 1.0
    bins = np.bincount(np.minimum((u * 100).astype(int), 99), minlength=100)
    exp = len(u) / 100
    chi2 = float(np.sum((bins - exp) ** 2 / exp))
    assert _chi2_sf(chi2, 99) > 1e-3, "chi2 %.1f" % chi2
    # genotype::sampleHaplotypeH0: allele 1 when u > P(allele 0)
    for p0 in (0.1, 0.5, 0.93):
        frac = float(np.mean(u > np.float32(p0)))
        sd = math.sqrt(p0 * (1 - p0) / len(u))
        assert abs(frac - (1 - p0)) < 5 * sd
