"""Pins of the concordance oracle (oracle/concordance_ref.py) against what the
mathematics and hand-computed cases fix, on the CPU:
* stats2D's Welford correlation against numpy's corrcoef, and its degenerate
  cases (basic_stats.h:110-115);
* getTruth / getMostLikely / calibration and frequency bins on hand-computed
  vectors (call_set_header.h:162-259);
* the counters and the non-reference discordance on a hand-built table
  (call_set_writing.cpp:46-55);
* the allele flip by mirror invariance;
* per-site r2 against corrcoef of the validated pairs.
Also the host-side r2 of the library (gh_conc_r2, no GPU) against the oracle."""
import math

import numpy as np
import pytest

import concordance_ref as cr
from conc_cases import conc_case

f32 = np.float32


def test_stats2d_matches_corrcoef():
    rng = np.random.default_rng(3)
    for n in (2, 3, 17, 500):
        x = rng.random(n) * 2
        y = rng.integers(0, 3, n).astype(float)
        if np.all(y == y[0]):
            y[0] = (y[0] + 1) % 3
        s = cr.Stats2D()
        for a, b in zip(x, y):
            s.push(a, b)
        want = np.corrcoef(x, y)[0, 1]
        assert abs(s.corr() - want) < 1e-12


def test_stats2d_degenerate():
    s = cr.Stats2D()
    assert s.corr() == 0.0                       # n = 0
    s.push(0.3, 1.0)
    assert math.isnan(s.corr())                  # n = 1: 0 / 0
    c = cr.Stats2D()
    for y in (0, 1, 2, 1):
        c.push(0.7, y)                           # constant x
    assert c.corr() == 0.0
    d = cr.Stats2D()
    for x in (0.1, 0.9, 1.7):
        d.push(x, 1)                             # constant y
    assert d.corr() == 0.0


def test_get_truth_hand_cases():
    u = cr.unphred()
    gl = lambda a, b, c: [u[a], u[b], u[c]]
    # PL 0,30,60: p0 = 1 / 1.001001 = 0.999
    assert cr.get_truth(gl(0, 30, 60), 2, 0.99, 0, 0) == 0
    assert cr.get_truth(gl(0, 30, 60), 2, 0.9999, 0, 0) == -4
    assert cr.get_truth(gl(40, 0, 40), 2, 0.9999, 0, 0) == -4       # 1 / 1.0002 < 0.9999
    assert cr.get_truth(gl(50, 0, 50), 2, 0.9999, 0, 0) == 1
    assert cr.get_truth(gl(60, 60, 0), 2, 0.9999, 0, 0) == 2
    assert cr.get_truth(gl(0, 0, 60), 2, 0.4, 0, 0) == -1          # tie between the two best
    assert cr.get_truth([f32(-1)] * 3, 2, 0.9, 0, 0) == -3
    assert cr.get_truth(gl(0, 30, 60), 2, 0.9, 3, 5) == -2          # DP 3 < 5
    assert cr.get_truth(gl(0, 30, 60), 2, 0.9, cr.BCF_INT32_MISSING, 5) == -5
    assert cr.get_truth(gl(0, 30, 60), 2, 0.9, 5, 5) == 0
    assert cr.get_truth([u[0], u[20], u[20]], 1, 0.99, 0, 0) == 0   # haploid: 1 / 1.01
    assert cr.get_truth([u[0], u[10], u[10]], 1, 0.99, 0, 0) == -4  # 1 / 1.1 = 0.909
    assert cr.get_truth([u[7], u[7], u[7]], 1, 0.4, 0, 0) == -1


def test_get_most_likely_hand_cases():
    g = lambda *v: [f32(x) for x in v]
    assert cr.get_most_likely(g(0.2, 0.5, 0.3), 2, 0.0) == 1
    assert cr.get_most_likely(g(0.2, 0.3, 0.5), 2, 0.0) == 2
    assert cr.get_most_likely(g(0.5, 0.5, 0.0), 2, 0.0) == 0        # diploid tie: 0
    assert cr.get_most_likely(g(0.5, 0.5, 0.5), 1, 0.0) == -1       # haploid tie: -1
    assert cr.get_most_likely(g(0.2, 0.5, 0.3), 2, 0.9) == -4
    assert cr.get_most_likely(g(0.95, 0.05, 0.0), 2, 0.9) == 0
    assert cr.get_most_likely(g(1.0000002, 0, 0), 2, 0.0) == -3    # above 1 + 1e-7
    assert cr.get_most_likely(g(1.0, 0, 0), 2, 0.0) == 0
    assert cr.get_most_likely(g(-0.001, 0.5, 0.5), 2, 0.0) == -3
    assert cr.calibration_bin(g(0.999, 0.001, 0)) == 98             # trunc(0.999 * 99) = 98
    assert cr.calibration_bin(g(1.0, 0, 0)) == 99
    assert cr.calibration_bin(g(0.34, 0.33, 0.33)) == 33
    bins = [0, 0.01, 0.05, 0.5]
    assert cr.frequency_bin(f32(0.01), bins) == 0                   # right edge is inclusive
    assert cr.frequency_bin(f32(0.0), bins) == -1                   # left edge excluded
    assert cr.frequency_bin(f32(0.3), bins) == 2


def _gt_case(pairs, P=2):
    """One site per (truth, estimate) pair, one sample: GT truth and GT estimates."""
    S = len(pairs)
    return {"n_sites": S, "n_samples": 1, "af": np.full(S, 0.2, np.float32), "type": np.ones(S, np.uint8),
            "truth_gt": np.array([[t] for t, _ in pairs], np.int8), "esti_gt": np.array([[e] for _, e in pairs], np.int8)}


def test_counters_and_nrd_hand_table():
    pairs = [(0, 0)] * 7 + [(0, 1)] * 2 + [(1, 1)] * 5 + [(1, 0)] * 1 + [(1, 2)] * 1 + [(2, 2)] * 3 + [(2, 1)] * 2
    inp = _gt_case(pairs)
    prm = {"ploidy": 2, "bins": [0, 0.5], "min_val_gl": 0.9, "min_tar_gp": 0.0, "truth_from_gt": 1, "esti_from_gt": 1}
    out = cr.concordance(prm, inp)
    cnt = out["spl"][0][0].cnt
    tab = cnt[:9].reshape(3, 3)
    assert tab.tolist() == [[7, 2, 0], [1, 5, 1], [0, 2, 3]]
    assert cnt[9:12].tolist() == [9, 7, 5]
    # NRD = mismatches (2 + 2 + 2) / (mismatches + non-ref matches (5 + 3))
    assert abs(cr.nrd_percent(cnt, 2) - 600.0 / 14.0) < 1e-12
    assert np.array_equal(out["grp"][0][0].cnt, cnt)
    # best-guess r2 against corrcoef of the pairs
    x = np.array([e for _, e in pairs], float)
    y = np.array([t for t, _ in pairs], float)
    assert abs(out["spl"][0][0].gt.corr() - np.corrcoef(x, y)[0, 1]) < 1e-12


def _mirror(inp, P):
    m = dict(inp)
    m["af"] = (np.float32(1.0) - inp["af"]).astype(np.float32)
    m["truth_pl"] = inp["truth_pl"][:, :, ::-1].copy()
    tg = inp["truth_gt"].astype(np.int16)
    m["truth_gt"] = np.where(tg < 0, -1, P - tg).astype(np.int8)
    m["esti_gp"] = inp["esti_gp"][:, :, ::-1].copy()
    m["esti_ds"] = (np.float32(P) - inp["esti_ds"]).astype(np.float32)
    eg = inp["esti_gt"].astype(np.int16)
    m["esti_gt"] = np.where(eg < 0, -1, P - eg).astype(np.int8)
    return m


@pytest.mark.parametrize("truth_gt", [0, 1])
def test_flip_mirror_invariance(truth_gt):
    """Swapping REF and ALT everywhere (af -> 1 - af, PL / GP reversed,
    DS -> 2 - DS) leaves every metric unchanged except at af = 0.5 exactly: the
    tool flips alleles at af > 0.5 so both inputs are classified alike."""
    c = conc_case(120, 12, seed=5)
    # af away from 0.5 and from 0 / 1 (1 - af must not round to 1 in float)
    ok = ~np.isnan(c["af"]) & (np.abs(c["af"] - 0.5) > 1e-6) & (c["af"] > 1e-6) & (c["af"] < 1.0 - 1e-6)
    inp = {k: (v[ok] if isinstance(v, np.ndarray) else v) for k, v in c.items()}
    inp.update(n_sites=int(ok.sum()), n_samples=12)
    prm = {"ploidy": 2, "bins": [0, 0.01, 0.05, 0.2, 0.5], "min_val_gl": 0.99, "min_tar_gp": 0.0,
           "truth_from_gt": truth_gt}
    a = cr.concordance(prm, inp)
    b = cr.concordance(prm, _mirror(inp, 2))
    for i in range(12):
        for t in range(2):
            assert np.array_equal(a["spl"][i][t].cnt, b["spl"][i][t].cnt)
            # DS -> 2 - DS -> 2 - (2 - DS) is exact only to float rounding of DS
            assert abs(a["spl"][i][t].ds.r2() - b["spl"][i][t].ds.r2()) < 1e-6 or (
                math.isnan(a["spl"][i][t].ds.r2()) and math.isnan(b["spl"][i][t].ds.r2()))
    assert np.array_equal(a["cal"], b["cal"])
    assert np.array_equal(a["rejected"], b["rejected"])


def test_mirror_without_flip_changes_metrics():
    """The invariance above is not vacuous: with use_alt_af (no flip) the
    mirrored input gives other counts."""
    c = conc_case(120, 12, seed=5)
    ok = ~np.isnan(c["af"]) & (c["af"] > 0.0) & (c["af"] < 1.0)
    inp = {k: (v[ok] if isinstance(v, np.ndarray) else v) for k, v in c.items()}
    inp.update(n_sites=int(ok.sum()), n_samples=12)
    prm = {"ploidy": 2, "bins": [0, 0.01, 0.05, 0.2, 0.5, 1.0], "min_val_gl": 0.99, "min_tar_gp": 0.0,
           "use_alt_af": 1}
    a = cr.concordance(prm, inp)
    b = cr.concordance(prm, _mirror(inp, 2))
    assert not all(np.array_equal(a["spl"][i][0].cnt, b["spl"][i][0].cnt) for i in range(12))


def test_site_r2_is_corrcoef_of_validated_pairs():
    c = conc_case(60, 30, seed=9)
    inp = dict(c, n_sites=60, n_samples=30)
    prm = {"ploidy": 2, "bins": [0, 0.5], "min_val_gl": 0.9, "min_tar_gp": 0.0}
    out = cr.concordance(prm, inp)
    # recompute the validated pairs of site s independently (no flip: af <= 0.5 sites only)
    checked = 0
    for s in range(60):
        if out["site_tag"][s] < 0 or c["af"][s] > 0.5 or out["site_r2"][s].n < 3:
            continue
        xs, ys = [], []
        for i in range(30):
            pl = c["truth_pl"][s, i]
            if np.any(pl < 0) or np.isnan(c["esti_gp"][s, i]).any() or c["esti_gt"][s, i] < 0:
                continue
            t = cr.get_truth([cr.unphred()[min(int(v), 255)] for v in pl], 2, 0.9, 0, 0)
            e = cr.get_most_likely([f32(v) for v in c["esti_gp"][s, i]], 2, 0.0)
            if t >= 0 and e >= 0:
                xs.append(float(c["esti_ds"][s, i]))
                ys.append(float(t))
        xs, ys = np.array(xs), np.array(ys)
        if np.ptp(xs) == 0 or np.ptp(ys) == 0:
            continue
        assert abs(out["site_r2"][s].corr() - np.corrcoef(xs, ys)[0, 1]) < 1e-12
        checked += 1
    assert checked > 10


def test_library_r2_matches_stats2d():
    """gh_conc_r2 (host code of the library) from sums against the Welford
    accumulator, including the degenerate cases."""
    from glimpse_amd import concordance as gc
    from glimpse_amd import hip
    L = gc._lib()
    rng = np.random.default_rng(1)
    for n in (0, 1, 2, 5, 1000):
        x = rng.random(n).astype(np.float32).astype(float) * 2
        y = rng.integers(0, 3, n).astype(float)
        s = cr.Stats2D()
        for a, b in zip(x, y):
            s.push(a, b)
        xc = int(n > 0 and np.ptp(x) == 0)
        yc = int(n > 0 and np.ptp(y) == 0)
        got = L.gh_conc_r2(n, x.sum(), (x * x).sum(), (x * y).sum(), y.sum(), (y * y).sum(), xc, yc)
        want = s.r2()
        if math.isnan(want):
            assert math.isnan(got)
        else:
            assert abs(got - want) < 1e-12
    # constant x: 0 whatever the sums say
    assert L.gh_conc_r2(4, 2.8, 1.96, 2.8, 4.0, 6.0, 1, 0) == 0.0
    assert hip.lib() is L
