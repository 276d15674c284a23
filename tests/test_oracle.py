"""Pins the CPU oracle against known answers and brute force (no GPU).

* std::mt19937: the C++ standard's known answer (10000th output of the
  default-seeded engine is 4123659995) -- the engine the reference's global
  rng wraps (common/src/utils/random_number.h:35-60).
* Philox4x32-10: Random123 known-answer vectors (Salmon et al., SC'11).
* unphred table: 10^(-i/10) with 7 significant digits (otools.h:97-98).
* Haploid forward-backward: a dense float64 textbook forward-backward of the
  same Li-Stephens model (stay 1-t plus switch t/K, emission by copied
  allele, own-site emission divided out, mixed with the imputation error and
  the site likelihood) on tiny panels.
"""
import numpy as np
import pytest

import orc
from glimpse_amd import synth


def test_mt19937_known_answer():
    assert orc.mt19937_nth(5489, 10000) == 4123659995


@pytest.mark.parametrize("ctr,key,out", [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
])
def test_philox_known_answer(ctr, key, out):
    assert [int(x) for x in orc.philox4x32(ctr, key)] == out


def test_unphred_table():
    vals = np.array([orc.unphred(i) for i in range(256)], np.float32)
    want = np.array([float("%.6e" % (10.0 ** (-i / 10.0))) for i in range(256)], np.float32)
    assert np.array_equal(vals, want)
    assert vals[0] == np.float32(1.0) and vals[1] == np.float32(7.943282e-01)
    assert vals[255] == np.float32(3.162278e-26)
    assert np.all(np.diff(vals) < 0)


def _dense_fb(hap_rows, poly, cm, HL, flat, lq, major, N, err=1e-12, ne=100000):
    """float64 forward-backward over the polymorphic sites; returns HP[2L]."""
    K = hap_rows.shape[0]
    L = HL.size // 2
    ee, ed = 1.0 - np.float64(np.float32(err)), np.float64(np.float32(err))
    nrho = -0.04 * ne / N
    HP = np.zeros(2 * L)
    Lp = len(poly)
    E = np.ones((Lp, K))
    hasE = np.zeros(Lp, bool)
    for i, s in enumerate(poly):
        if not (flat[s] or lq[s]):
            p0 = HL[2 * s] * ee + HL[2 * s + 1] * ed
            p1 = HL[2 * s] * ed + HL[2 * s + 1] * ee
            e = np.array([p0, p1]) / (p0 + p1)
            E[i] = e[hap_rows[:, s]]
            hasE[i] = True
    tr = [np.clip(-np.expm1(nrho * (cm[poly[i]] - cm[poly[i - 1]])), 1e-7, 1 - 1e-7) for i in range(1, Lp)]
    F = np.zeros((Lp, K))
    B = np.zeros((Lp, K))
    F[0] = E[0] / K
    for i in range(1, Lp):
        t = tr[i - 1]
        F[i] = ((1 - t) * F[i - 1] + t / K * F[i - 1].sum()) * E[i]
        F[i] /= F[i].sum()
    B[-1] = 1.0
    for i in range(Lp - 2, -1, -1):
        t = tr[i]
        x = E[i + 1] * B[i + 1]
        B[i] = (1 - t) * x + t / K * x.sum()
        B[i] /= B[i].sum()
    for i, s in enumerate(poly):
        post = F[i] * B[i]
        ph = np.array([post[hap_rows[:, s] == 0].sum(), post[hap_rows[:, s] == 1].sum()])
        if hasE[i]:
            p0 = HL[2 * s] * ee + HL[2 * s + 1] * ed
            p1 = HL[2 * s] * ed + HL[2 * s + 1] * ee
            ph = ph / (np.array([p0, p1]) / (p0 + p1))
        po = np.array([ph[0] * ee + ph[1] * ed, ph[0] * ed + ph[1] * ee])
        if not flat[s]:
            po = po * HL[2 * s: 2 * s + 2]
        HP[2 * s: 2 * s + 2] = po / po.sum()
    for s in range(L):
        if s in set(poly):
            continue
        po = np.array([ed, ed])
        po[major[s]] = ee
        if not flat[s]:
            po = po * HL[2 * s: 2 * s + 2]
        HP[2 * s: 2 * s + 2] = po / po.sum()
    return HP


@pytest.mark.parametrize("K,seed", [(1, 1), (3, 2), (8, 3), (13, 4), (30, 5)])
def test_forward_backward_matches_dense_float64(K, seed):
    N, L = 2200, 60
    P = synth.make_panel(N, L, 1, length_mb=0.05, rare_frac=0.4, seed=seed, keep_dense=True)
    gl, flat = synth.make_reads(P["tgt_haps"], 2.0, seed=seed)
    P["lq"] = (np.arange(P["L"]) % 9 == 4).astype(np.uint8)
    OP = orc.Panel(P)
    rng = np.random.default_rng(seed)
    states = np.sort(rng.choice(N, K, replace=False))
    HL = orc.hap_likelihoods(gl[0], flat[0])
    HP = OP.impute_hap(states, HL, flat[0])
    cs = OP.condset_from_states(states)
    fc, mj = OP.site_class()
    want = _dense_fb(P["dense"][states], cs["poly"], P["cm"], HL.astype(np.float64), flat[0], P["lq"], mj, N)
    assert np.abs(HP - want).max() < 2e-5
    assert np.allclose(HP[0::2] + HP[1::2], 1.0, atol=1e-6)


def test_hap_likelihoods_closed_form():
    """initHaplotypeLikelihoods: marginal HL0 = g0 + g1/2, HL1 = g1/2 + g2 from
    normalised unphred GLs; makeHaplotypeLikelihoods conditions on the other
    allele; flat sites give (0.5, 0.5) in init mode; min_gl clamps."""
    gl = np.array([[0, 10, 20], [30, 0, 30], [40, 20, 0], [0, 0, 0], [0, 255, 255]], np.uint8)
    flat = np.array([0, 0, 0, 1, 0], np.uint8)
    HL = orc.hap_likelihoods(gl, flat).reshape(-1, 2)
    for l in range(5):
        if flat[l]:
            assert np.allclose(HL[l], 0.5)
            continue
        g = 10.0 ** (-gl[l].astype(np.float64) / 10.0)
        g /= g.sum()
        want = np.array([g[0] + g[1] / 2, g[1] / 2 + g[2]])
        want = np.maximum(want, 1e-10)
        assert np.allclose(HL[l], want / want.sum(), rtol=1e-5, atol=1e-9)
    other = np.array([1, 0, 1, 0, 1], np.uint8)
    HC = orc.hap_likelihoods(gl, flat, other, HL=HL.ravel()).reshape(-1, 2)
    for l in range(5):
        if flat[l]:
            continue
        g = 10.0 ** (-gl[l].astype(np.float64) / 10.0)
        c = other[l]
        want = np.array([g[c], g[c + 1]]) / (g[c] + g[c + 1])
        if want[0] < 1e-10:
            want = np.array([1e-10, 1 - 1e-10])
        if want[1] < 1e-10:
            want = np.array([1 - 1e-10, 1e-10])
        assert np.allclose(HC[l], want, rtol=1e-5, atol=1e-9)


def test_reference_mode_is_deterministic_and_accurate():
    P = synth.make_panel(500, 2000, 2, length_mb=0.4, seed=5)
    gl, flat = synth.make_reads(P["tgt_haps"], 1.0, seed=6)
    OP = orc.Panel(P)
    prm = orc.default_params(burnin=2, main_iter=3, kpbwt=200, kinit=300)
    a = OP.phase_run(gl, flat, prm)
    b = OP.phase_run(gl, flat, prm)
    assert np.array_equal(a["gp0"], b["gp0"]) and np.array_equal(a["gp1"], b["gp1"])
    truth = synth.truth_genotypes(P["tgt_haps"])
    ds = a["gp1"] + 2 * np.clip(1 - a["gp0"] - a["gp1"], 0, 1)
    r = np.corrcoef(ds.ravel(), truth.ravel())[0, 1]
    assert r * r > 0.6
