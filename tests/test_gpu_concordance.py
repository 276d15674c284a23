"""GPU parity: gh_concordance against the concordance oracle
(oracle/concordance_ref.py, restating call_set_reading.cpp:476-760).
Counters (tables, usable truths, filtered, calibration, rejections, site
groups) are bit-exact; correlations agree to double rounding (the oracle sums
with Welford updates, the GPU with fixed-order double sums)."""
import math

import numpy as np
import pytest

import concordance_ref as cr
from conc_cases import conc_case

pytestmark = pytest.mark.gpu

BINS = [0, 0.001, 0.005, 0.01, 0.05, 0.1, 0.2, 0.5]


def _close(a, b, tol=1e-9):
    if math.isnan(a) or math.isnan(b):
        return math.isnan(a) and math.isnan(b)
    return abs(a - b) <= tol


def _run_both(c, S, N, **kw):
    from glimpse_amd import concordance as gc
    P = c["ploidy"]
    mode_truth_gt = kw.pop("truth_gt", False)
    mode_esti_gt = kw.pop("esti_gt", False)
    args = dict(ploidy=P, af=c["af"][:S], type_=c["type"][:S])
    if mode_truth_gt:
        args["truth_gt"] = c["truth_gt"][:S, :N]
    else:
        args["truth_pl"] = c["truth_pl"][:S, :N]
    if kw.get("min_val_dp", 0) > 0:
        args["truth_dp"] = c["truth_dp"][:S, :N]
    if mode_esti_gt:
        args["esti_gt"] = c["esti_gt"][:S, :N]
    else:
        args.update(esti_gp=c["esti_gp"][:S, :N], esti_ds=c["esti_ds"][:S, :N], esti_gt=c["esti_gt"][:S, :N])
    args.update(kw)
    got = gc.run(**args)
    prm = {"ploidy": P, "min_val_gl": kw.get("min_val_gl", 0.9999), "min_val_dp": kw.get("min_val_dp", 0),
           "min_tar_gp": kw.get("min_tar_gp", 0.0), "use_alt_af": int(kw.get("use_alt_af", False)),
           "truth_from_gt": int(mode_truth_gt), "esti_from_gt": int(mode_esti_gt)}
    if kw.get("site_group") is not None:
        prm.update(site_group=kw["site_group"], n_groups=kw["n_groups"], bins=None)
    else:
        prm["bins"] = kw.get("bins", gc.DEFAULT_BINS)
    inp = {k: v for k, v in args.items() if k in ("truth_pl", "truth_gt", "truth_dp", "esti_gp", "esti_ds", "esti_gt")}
    inp.update(af=args["af"], type=args["type_"], n_sites=S, n_samples=N)
    want = cr.concordance(prm, inp)
    return got, want


def _compare(got, want, P):
    from glimpse_amd import concordance as gc
    N = got["spl_cnt"].shape[0]
    G = got["grp_cnt"].shape[0]
    for i in range(N):
        for t in range(2):
            rec = want["spl"][i][t]
            assert np.array_equal(got["spl_cnt"][i, t].astype(np.int64), rec.cnt), "sample %d type %d" % (i, t)
            assert _close(gc.r2_of(got["spl_cnt"][i, t], got["spl_sum"][i, t], "ds"), rec.ds.r2())
            assert _close(gc.r2_of(got["spl_cnt"][i, t], got["spl_sum"][i, t], "gt"), rec.gt.r2())
    for g in range(G):
        for t in range(2):
            rec = want["grp"][g][t]
            assert np.array_equal(got["grp_cnt"][g, t].astype(np.int64), rec.cnt), "group %d type %d" % (g, t)
            assert _close(gc.r2_of(got["grp_cnt"][g, t], got["grp_sum"][g, t], "ds"), rec.ds.r2())
            assert _close(gc.r2_of(got["grp_cnt"][g, t], got["grp_sum"][g, t], "gt"), rec.gt.r2())
            n = int(got["grp_cnt"][g, t][:9].sum())
            assert n == rec.freq.n
            if n:
                assert abs(got["grp_sum"][g, t, 5] / n - rec.freq.mean()) < 1e-12
    assert np.array_equal(got["cal_cnt"].astype(np.int64), want["cal"])
    assert np.array_equal(got["rejected"].astype(np.int64), want["rejected"])
    assert np.array_equal(got["site_tag"], want["site_tag"])
    for s, st in enumerate(want["site_r2"]):
        n = int(got["site_cnt"][s][:9].sum())
        assert n == st.n
        assert _close(gc.r2_of(got["site_cnt"][s], got["site_sum"][s], "ds"), st.r2())


CASES = [
    # name, ploidy, seed, S, N, kwargs
    ("pl_gp", 2, 1, 400, 70, {}),
    ("pl_gp_dp_filter", 2, 2, 300, 50, {"min_val_dp": 8, "min_tar_gp": 0.9, "min_val_gl": 0.99}),
    ("gt_gt", 2, 3, 300, 64, {"truth_gt": True, "esti_gt": True}),
    ("gt_truth_gp", 2, 4, 250, 33, {"truth_gt": True, "min_tar_gp": 0.6}),
    ("haploid", 1, 5, 300, 40, {"min_val_gl": 0.99}),
    ("alt_af", 2, 6, 300, 45, {"use_alt_af": True, "bins": [0, 0.01, 0.1, 0.5, 1.0]}),
    ("many_samples", 2, 7, 60, 700, {}),
]


@pytest.mark.parametrize("name,P,seed,S,N,kw", CASES, ids=[c[0] for c in CASES])
def test_concordance_matches_oracle(name, P, seed, S, N, kw):
    c = conc_case(S, N, ploidy=P, seed=seed)
    got, want = _run_both(c, S, N, **dict(kw))
    _compare(got, want, P)
    assert int(got["spl_cnt"][:, :, :9].sum()) > S   # genotypes were validated


def test_concordance_site_groups():
    S, N = 300, 40
    c = conc_case(S, N, seed=8)
    rng = np.random.default_rng(8)
    grp = rng.integers(-1, 6, S).astype(np.int32)
    grp[rng.random(S) < 0.05] = 9   # beyond n_groups: not used
    got, want = _run_both(c, S, N, site_group=grp, n_groups=6)
    _compare(got, want, 2)


def test_concordance_large_reductions_agree():
    """At 20,000 sites x 400 samples: per-sample, per-group and per-site
    records count the same genotypes; the calibration table holds every
    estimated genotype; two runs are bit-identical (fixed reduction order)."""
    from glimpse_amd import concordance as gc
    S, N = 20000, 400
    c = conc_case(S, N, seed=21)
    kw = dict(ploidy=2, af=c["af"], type_=c["type"], truth_pl=c["truth_pl"], esti_gp=c["esti_gp"],
              esti_ds=c["esti_ds"], esti_gt=c["esti_gt"], bins=BINS)
    a = gc.run(**kw)
    b = gc.run(**kw)
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    spl = a["spl_cnt"].astype(np.int64).sum(axis=0)
    grp = a["grp_cnt"].astype(np.int64).sum(axis=0)
    tags = a["site_tag"]
    site = np.zeros((2, 13), np.int64)
    for t in range(2):
        site[t] = a["site_cnt"][(tags >= 0) & ((tags & 1) == t)].astype(np.int64).sum(axis=0)
    assert np.array_equal(spl, grp) and np.array_equal(spl, site)
    assert np.array_equal(a["cal_cnt"].astype(np.int64).sum(axis=0), spl[:, :9])
    assert np.isclose(a["spl_sum"][:, :, 0].sum(), a["grp_sum"][:, :, 0].sum(), rtol=1e-12)
    # oracle parity on a slice of the same input
    got, want = _run_both(c, 150, 60, bins=BINS)
    _compare(got, want, 2)
