"""CPU restatement of the concordance tool's genotype metrics -- test
infrastructure only (the checker of glimpse_amd's gh_concordance), never part
of the product path.

Follows, per site and sample, concordance/src/containers/call_set_reading.cpp:
476-760 (site filter, allele flip at af > 0.5, frequency bin, truth from PL or
GT with the DP rule, estimate from GP/DS or GT, the counters and the stats2D
pushes), the rules of call_set_header.h:162-259 (getTruth, getMostLikely,
getCalibrationBin, getFrequencyBin) and the Welford accumulators of
concordance/src/utils/basic_stats.h:32-125, and the derived columns of
call_set_writing.cpp:42-90 (non-reference discordance, r2 rounding).

Float32 arithmetic is reproduced with numpy float32 scalars where the
reference computes in float.  Pure Python loops: small inputs only.

One departure, for ploidy 1: the reference's NaN test of GP reads the third
value of each sample's GP record (call_set_reading.cpp:623), which for a
haploid record is the next sample's first value; here only the sample's own
values are tested (as the GPU path does).

How each part is pinned (tests/test_concordance_oracle.py): stats2D against
numpy's corrcoef (closed form) and its degenerate cases; getTruth and
getMostLikely on hand-computed probability vectors; the counters and NRD
against a hand-built confusion table; the allele flip by mirror invariance
(swapping REF/ALT of every input and af -> 1 - af gives identical metrics).
"""
import math

import numpy as np

import orc

f32 = np.float32
N_BIN_CAL = 100
BCF_INT32_MISSING = -2147483648
BCF_INT32_VECTOR_END = -2147483647

_UNPH = None


def unphred():
    global _UNPH
    if _UNPH is None:
        _UNPH = [f32(orc.unphred(i)) for i in range(256)]
    return _UNPH


class Stats2D:
    """basic_stats.h:32-125: Welford updates in double."""

    def __init__(self):
        self.n = 0
        self.oMx = self.nMx = 0.0
        self.oSx = self.nSx = 0.0
        self.oMy = self.nMy = 0.0
        self.oSy = self.nSy = 0.0
        self.oC = self.nC = 0.0

    def push(self, x, y):
        x = float(x)
        y = float(y)
        self.n += 1
        if self.n == 1:
            self.oMx = self.nMx = x
            self.oMy = self.nMy = y
            self.oSx = self.oSy = 0.0
            self.oC = self.nC = 0.0
        else:
            self.nMy = self.oMy + (y - self.oMy) / self.n
            self.nC = self.oC + (x - self.oMx) * (y - self.nMy)
            self.nMx = self.oMx + (x - self.oMx) / self.n
            self.nSx = self.oSx + (x - self.oMx) * (x - self.nMx)
            self.nSy = self.oSy + (y - self.oMy) * (y - self.nMy)
            self.oC = self.nC
            self.oMx = self.nMx
            self.oSx = self.nSx
            self.oMy = self.nMy
            self.oSy = self.nSy

    def corr(self):
        n = self.n
        sdx = math.sqrt(self.nSx / (n - 1)) if (n > 1 and self.nSx > 0) else 1e-10
        sdy = math.sqrt(self.nSy / (n - 1)) if (n > 1 and self.nSy > 0) else 1e-10
        if n == 0:
            return 0.0
        den = (n - 1) * sdx * sdy
        if den == 0.0:
            return float("nan") if self.nC == 0.0 else math.copysign(float("inf"), self.nC)
        return self.nC / den

    def r2(self):
        r = self.corr()
        return r * r


class Stats1D:
    def __init__(self):
        self.n = 0
        self.s = 0.0

    def push(self, x):
        self.n += 1
        self.s += float(x)

    def mean(self):
        return self.s / self.n if self.n else 0.0


def get_truth(gl, P, T, dp, D):
    """getTruth (call_set_header.h:175-205); gl: three float32 values."""
    if D > 0 and dp == BCF_INT32_MISSING:
        return -5
    if dp < D:
        return -2
    T = f32(T)
    if P == 1:
        g0, g1 = gl[0], gl[1]
        if g0 < 0 or g1 < 0:
            return -3
        sc = f32(1.0) / f32(g0 + g1)
        p0, p1 = f32(g0 * sc), f32(g1 * sc)
        if p0 < T and p1 < T:
            return -4
        if p0 > p1:
            return 0
        if p1 > p0:
            return 1
        return -1
    g0, g1, g2 = gl
    if g0 < 0 or g1 < 0 or g2 < 0:
        return -3
    sc = f32(1.0 / float(f32(f32(g0 + g1) + g2)))
    p0, p1, p2 = f32(g0 * sc), f32(g1 * sc), f32(g2 * sc)
    if p0 < T and p1 < T and p2 < T:
        return -4
    if p0 > p1 and p0 > p2:
        return 0
    if p1 > p0 and p1 > p2:
        return 1
    if p2 > p0 and p2 > p1:
        return 2
    return -1


def get_most_likely(gp, P, gp_filter):
    """getMostLikely (call_set_header.h:214-246)."""
    hi = 1.0 + 1e-7
    g0, g1, g2 = gp
    if g0 < 0 or float(g0) > hi or g1 < 0 or float(g1) > hi:
        return -3
    if P == 1:
        if max(g0, g1) < f32(gp_filter):
            return -4
        if g0 > g1:
            return 0
        if g1 > g0:
            return 1
        return -1
    if g2 < 0 or float(g2) > hi:
        return -3
    if max(g0, g1, g2) < f32(gp_filter):
        return -4
    if g0 > g1 and g0 > g2:
        return 0
    if g1 > g0 and g1 > g2:
        return 1
    if g2 > g0 and g2 > g1:
        return 2
    return 0


def calibration_bin(gp):
    return int(math.trunc(float(f32(max(gp[0], max(gp[1], gp[2])) * f32(N_BIN_CAL - 1)))))


def frequency_bin(maf, bins):
    for b in range(1, len(bins)):
        if float(maf) > bins[b - 1] and float(maf) <= bins[b]:
            return b - 1
    return -1


class Record:
    """Counters and accumulators of one (sample | group | site, type)."""

    def __init__(self):
        self.cnt = np.zeros(13, np.int64)
        self.ds = Stats2D()
        self.gt = Stats2D()
        self.freq = Stats1D()


def concordance(prm, inp):
    """prm / inp: dicts with the fields of gh_conc_params / gh_conc_input
    (numpy arrays).  Returns dict with records per sample [N][2], group
    [G][2], calibration counts [100][2][9], rejected [11], per-site r2
    records [S] (Stats2D of DS against truth) and site tags."""
    P = prm["ploidy"]
    S, N = inp["n_sites"], inp["n_samples"]
    bins = prm.get("bins")
    n_bins = len(bins) - 1 if bins is not None and len(bins) else 0
    G = n_bins if n_bins > 0 else prm["n_groups"]
    T, D, gpf = prm["min_val_gl"], prm.get("min_val_dp", 0), prm["min_tar_gp"]
    alt = prm.get("use_alt_af", 0)
    unph = unphred()
    spl = [[Record(), Record()] for _ in range(N)]
    grp = [[Record(), Record()] for _ in range(G)]
    cal = np.zeros((N_BIN_CAL, 2, 9), np.int64)
    rej = np.zeros(11, np.int64)
    site_r2 = [Stats2D() for _ in range(S)]
    site_tag = np.full(S, -1, np.int32)
    for s in range(S):
        af = f32(inp["af"][s])
        ty = int(inp["type"][s])
        if math.isnan(af) or ty not in (1, 2):
            rej[9] += 1
            continue
        flip = False if alt else bool(af > f32(0.5))
        maf = af if alt else min(af, f32(f32(1.0) - af))
        if n_bins > 0:
            g = frequency_bin(maf, bins)
        else:
            g = int(prm["site_group"][s])
            if g >= G:
                g = -1
        if g < 0:
            rej[10] += 1
            continue
        t = 1 if ty == 2 else 0
        site_tag[s] = 2 * g + t
        for i in range(N):
            # truth
            dp = int(inp["truth_dp"][s, i]) if D > 0 else 0
            if prm.get("truth_from_gt"):
                gtv = int(inp["truth_gt"][s, i])
                if gtv < 0:
                    gl = [f32(-1.0)] * 3
                else:
                    gl = [f32(gtv == 0), f32(gtv == 1), f32(gtv == P)]
            else:
                pl = [int(v) for v in inp["truth_pl"][s, i]]
                if any(v < 0 for v in pl):
                    gl = [f32(-1.0)] * 3
                else:
                    gl = [unph[min(pl[0], 255)], unph[min(pl[1], 255)], unph[min(pl[P], 255)]]
            if P == 1:
                gl[2] = gl[1]
            if flip:
                gl[0], gl[2] = gl[2], gl[0]
                if P == 1:
                    gl[1] = gl[2]
            tg = get_truth(gl, P, T, dp, D)
            if tg < 0:
                rej[-tg - 1] += 1
                continue
            # estimate
            if prm.get("esti_from_gt"):
                gtv = int(inp["esti_gt"][s, i])
                miss = gtv < 0
                ds = f32(gtv)
                gp = [f32(gtv == 0), f32(gtv == 1), f32(gtv == P)]
            else:
                miss = inp.get("esti_gt") is not None and int(inp["esti_gt"][s, i]) < 0
                rec = inp["esti_gp"][s, i]
                gp = [f32(rec[0]), f32(rec[1]), f32(rec[P])]
                miss = miss or any(math.isnan(v) for v in gp)
                ds = f32(inp["esti_ds"][s, i])
            if miss:
                es = -3
            else:
                if flip:
                    gp[0], gp[2] = gp[2], gp[0]
                    if P == 1:
                        gp[1] = gp[2]
                    ds = f32(f32(P) - ds)
                es = get_most_likely(gp, P, gpf)
            for R in (spl[i][t], grp[g][t]):
                R.cnt[9 + tg] += 1
            if es >= 0:
                cb = calibration_bin(gp)
                for R in (spl[i][t], grp[g][t]):
                    R.cnt[3 * tg + es] += 1
                    R.ds.push(ds, tg)
                    R.gt.push(es, tg)
                grp[g][t].freq.push(maf)
                cal[cb, t, 3 * tg + es] += 1
                site_r2[s].push(ds, tg)
            else:
                if es == -4:
                    spl[i][t].cnt[12] += 1
                    grp[g][t].cnt[12] += 1
                rej[5 + (-es - 1)] += 1
    return {"spl": spl, "grp": grp, "cal": cal, "rejected": rej, "site_r2": site_r2, "site_tag": site_tag}


def nrd_percent(cnt, P):
    """Non-reference discordance of a record (call_set_writing.cpp:46-55):
    mismatches at every truth over mismatches + matches at non-reference
    truths, in percent."""
    tab = np.asarray(cnt[:9]).reshape(3, 3)
    err = [int(tab[t].sum() - tab[t, t]) for t in range(3)]
    tot = [int(tab[t].sum()) for t in range(3)]
    mm = err[0] + sum(err[1:P + 1])
    m = sum(tot[p] - err[p] for p in range(1, P + 1))
    return mm * 100.0 / (m + mm) if (m + mm) > 0 else 0.0


def rounded_r2(stats):
    """The writer's r2 column: round(corr^2 * 1e6) / 1e6."""
    r = stats.corr()
    return round(r * r * 1000000.0) / 1000000.0 if not math.isnan(r) else float("nan")
