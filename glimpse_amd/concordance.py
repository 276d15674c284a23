"""Host mirror of the concordance tool (concordance/src/containers/call_set_*.cpp)
over gh_concordance: the per-site / per-sample classification and counting run
on the GPU; this module holds the call_set-like front end and the derived
tables of call_set_writing.cpp (matches, mismatches, rates, non-reference
discordance, r2 of best-guess genotypes and of dosages, calibration).

Inputs are arrays already parsed from the truth, estimate and frequency
files: the reference's htslib parsing (call_set_reading.cpp:398-475) is
outside this module.  No CPU fallback: every call goes to libglimpse_hip.so.
"""
import ctypes
import math

import numpy as np

from . import hip

vp = ctypes.c_void_p
NCNT, NSUM, NCAL, NREJ = 13, 5, 100, 11
REJECT_NAMES = ["truth_-1_tie", "truth_-2_low_dp", "truth_-3_missing", "truth_-4_uncertain", "truth_-5_no_dp",
                "esti_-1_tie", "esti_-2", "esti_-3_missing", "esti_-4_filtered", "site_no_af_or_type", "site_no_bin"]
# Default frequency bins of the concordance tool's examples (MAF bins).
DEFAULT_BINS = [0, 0.001, 0.002, 0.005, 0.01, 0.02, 0.05, 0.1, 0.2, 0.5]


class ConcParams(ctypes.Structure):
    _fields_ = [("ploidy", ctypes.c_int32), ("n_bins", ctypes.c_int32), ("bins", vp), ("n_groups", ctypes.c_int32),
                ("site_group", vp), ("min_val_gl", ctypes.c_float), ("min_val_dp", ctypes.c_int32),
                ("min_tar_gp", ctypes.c_float), ("use_alt_af", ctypes.c_int32), ("truth_from_gt", ctypes.c_int32),
                ("esti_from_gt", ctypes.c_int32)]


class ConcInput(ctypes.Structure):
    _fields_ = [("n_sites", ctypes.c_int32), ("n_samples", ctypes.c_int32), ("af", vp), ("type", vp),
                ("truth_pl", vp), ("truth_gt", vp), ("truth_dp", vp), ("esti_gp", vp), ("esti_ds", vp),
                ("esti_gt", vp)]


class ConcOutput(ctypes.Structure):
    _fields_ = [("spl_cnt", vp), ("spl_sum", vp), ("grp_cnt", vp), ("grp_sum", vp), ("cal_cnt", vp),
                ("rejected", vp), ("site_cnt", vp), ("site_sum", vp), ("site_tag", vp)]


_bound = False


def _lib():
    global _bound
    L = hip.lib()
    if not _bound:
        L.gh_concordance.argtypes = [ctypes.POINTER(ConcParams), ctypes.POINTER(ConcInput), ctypes.POINTER(ConcOutput)]
        L.gh_conc_r2.restype = ctypes.c_double
        L.gh_conc_r2.argtypes = [ctypes.c_uint64, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_double, ctypes.c_double, ctypes.c_int32, ctypes.c_int32]
        _bound = True
    return L


def _arr(a, dt):
    return None if a is None else np.ascontiguousarray(a, dt)


def run(ploidy, af, type_, esti_gp=None, esti_ds=None, esti_gt=None, truth_pl=None, truth_gt=None, truth_dp=None,
        bins=DEFAULT_BINS, site_group=None, n_groups=0, min_val_gl=0.9999, min_val_dp=0, min_tar_gp=0.0,
        use_alt_af=False, per_site=True):
    """gh_concordance over S sites x N samples.  Returns the raw records:
    spl_cnt [N][2][13], spl_sum [N][2][5], grp_cnt [G][2][13], grp_sum [G][2][6],
    cal_cnt [100][2][9], rejected [11] and, with per_site, site_cnt [S][13],
    site_sum [S][5], site_tag [S]."""
    af = _arr(af, np.float32)
    type_ = _arr(type_, np.uint8)
    S = af.shape[0]
    truth_from_gt = truth_pl is None
    esti_from_gt = esti_gp is None
    truth_pl, truth_gt, truth_dp = _arr(truth_pl, np.int32), _arr(truth_gt, np.int8), _arr(truth_dp, np.int32)
    esti_gp, esti_ds, esti_gt = _arr(esti_gp, np.float32), _arr(esti_ds, np.float32), _arr(esti_gt, np.int8)
    ref = truth_gt if truth_from_gt else truth_pl
    N = ref.shape[1]
    prm = ConcParams()
    prm.ploidy = ploidy
    if site_group is None:
        bins = np.ascontiguousarray(bins, np.float64)
        prm.n_bins = len(bins) - 1
        prm.bins = bins.ctypes.data_as(vp)
        G = prm.n_bins
    else:
        site_group = np.ascontiguousarray(site_group, np.int32)
        prm.n_groups = n_groups
        prm.site_group = site_group.ctypes.data_as(vp)
        G = n_groups
    prm.min_val_gl, prm.min_val_dp, prm.min_tar_gp = min_val_gl, min_val_dp, min_tar_gp
    prm.use_alt_af, prm.truth_from_gt, prm.esti_from_gt = int(use_alt_af), int(truth_from_gt), int(esti_from_gt)
    inp = ConcInput(S, N, hip._p(af), hip._p(type_), hip._p(truth_pl), hip._p(truth_gt), hip._p(truth_dp),
                    hip._p(esti_gp), hip._p(esti_ds), hip._p(esti_gt))
    out = {"spl_cnt": np.zeros((N, 2, NCNT), np.uint64), "spl_sum": np.zeros((N, 2, NSUM)),
           "grp_cnt": np.zeros((G, 2, NCNT), np.uint64), "grp_sum": np.zeros((G, 2, NSUM + 1)),
           "cal_cnt": np.zeros((NCAL, 2, 9), np.uint64), "rejected": np.zeros(NREJ, np.uint64)}
    if per_site:
        out.update(site_cnt=np.zeros((S, NCNT), np.uint32), site_sum=np.zeros((S, NSUM)),
                   site_tag=np.zeros(S, np.int32))
    o = ConcOutput(*[hip._p(out.get(k)) for k, _ in ConcOutput._fields_])
    hip._check(_lib().gh_concordance(ctypes.byref(prm), ctypes.byref(inp), ctypes.byref(o)))
    return out


def r2_of(cnt, sums, which="ds"):
    """Squared correlation (stats2D::corrXY^2) of one record: DS against the
    truth (which="ds") or the best-guess genotype against the truth ("gt")."""
    cnt = np.asarray(cnt, np.uint64).astype(np.int64)
    tab = cnt[:9].reshape(3, 3)
    n = int(tab.sum())
    rows = tab.sum(axis=1)
    sy = float(rows[1] + 2 * rows[2])
    syy = float(rows[1] + 4 * rows[2])
    y_const = int((rows > 0).sum() <= 1)
    if which == "gt":
        cols = tab.sum(axis=0)
        sx = float(cols[1] + 2 * cols[2])
        sxx = float(cols[1] + 4 * cols[2])
        sxy = float(tab[1, 1] + 2 * tab[1, 2] + 2 * tab[2, 1] + 4 * tab[2, 2])
        x_const = int((cols > 0).sum() <= 1)
    else:
        sx, sxx, sxy = float(sums[0]), float(sums[1]), float(sums[2])
        x_const = int(n > 0 and sums[3] == sums[4])
    return _lib().gh_conc_r2(n, sx, sxx, sxy, sy, syy, x_const, y_const)


def record_table(cnt, sums, ploidy):
    """The columns call_set_writing.cpp writes for one record (:42-90):
    val counts, filtered, matches, mismatches, mismatch rates (%), NRD (%),
    best-guess r2 and dosage r2 (both rounded to 1e-6)."""
    cnt = np.asarray(cnt, np.uint64).astype(np.int64)
    tab = cnt[:9].reshape(3, 3)
    tot = tab.sum(axis=1)
    err = tot - np.diag(tab)
    P = ploidy
    nrd_mm = int(err[0] + err[1:P + 1].sum())
    nrd_m = int((tot[1:P + 1] - err[1:P + 1]).sum())
    rg, rd = r2_of(cnt, sums, "gt"), r2_of(cnt, sums, "ds")
    rnd = lambda v: math.floor(v * 1000000.0 + 0.5) / 1000000.0 if not math.isnan(v) else float("nan")
    return {
        "val": [int(v) for v in cnt[9:9 + P + 1]],
        "filtered": int(cnt[12]),
        "matches": [int(tot[t] - err[t]) for t in range(P + 1)],
        "mismatches": [int(err[t]) for t in range(P + 1)],
        "mismatch_pct": [err[t] * 100.0 / tot[t] if tot[t] > 0 else 0.0 for t in range(P + 1)],
        "nrd_pct": nrd_mm * 100.0 / (nrd_m + nrd_mm) if (nrd_m + nrd_mm) > 0 else 0.0,
        "r2_gt": rnd(rg),
        "r2_ds": rnd(rd),
    }


def merge_types(cnt, sums):
    """The reference's "all variants" record from the SNP and indel ones."""
    cnt = np.asarray(cnt)
    sums = np.asarray(sums)
    c = cnt[..., 0, :] + cnt[..., 1, :]
    s = sums[..., 0, :].copy()
    s[..., :3] += sums[..., 1, :3]
    s[..., 3] = np.minimum(sums[..., 0, 3], sums[..., 1, 3])
    s[..., 4] = np.maximum(sums[..., 0, 4], sums[..., 1, 4])
    if s.shape[-1] > 5:
        s[..., 5] += sums[..., 1, 5]
    return c, s


def summary(out, ploidy):
    """Overall figures over every validated genotype of every type: dosage r2,
    best-guess r2, NRD and NRC (= 1 - NRD), genotype error rate."""
    c, s = merge_types(out["spl_cnt"], out["spl_sum"])
    tot_c = c.astype(np.int64).sum(axis=0)
    tot_s = np.array([s[:, 0].sum(), s[:, 1].sum(), s[:, 2].sum(), s[:, 3].min(), s[:, 4].max()])
    t = record_table(tot_c, tot_s, ploidy)
    tab = tot_c[:9].reshape(3, 3)
    n = int(tab.sum())
    return {"n": n, "r2_ds": r2_of(tot_c, tot_s, "ds"), "r2_gt": r2_of(tot_c, tot_s, "gt"),
            "nrd": t["nrd_pct"] / 100.0, "nrc": 1.0 - t["nrd_pct"] / 100.0,
            "error_rate": (n - int(np.trace(tab))) / n if n else 0.0}
