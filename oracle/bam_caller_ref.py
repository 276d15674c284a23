"""CPU restatement of the phase tool's BAM genotype-likelihood caller -- test
infrastructure only (the checker of gh_bam_call_gls), never part of the
product path.

Follows phase/src/io/genotype_bam_caller.cpp for SNP sites with the
"standard" model and indels left flat (the reference's default, without
--call-indels):
* read_bam (:30-59): reads on no contig, with an excluded flag or a mapping
  quality below the threshold are dropped; paired reads with an unmapped mate
  are dropped unless orphans are kept; with the orientation check, mates on
  another contig, on the same strand, or facing away are dropped; with the
  proper-pair check, reads not properly paired are dropped;
* the pileup (htslib's bam_plp_auto): every kept read whose alignment covers
  the site's position contributes the base and quality at that position when
  the position lies in an M, = or X operation;
* bcf_call_glfgen_min (:270-290): bases other than A, C, G, T and qualities
  below the base-quality threshold are dropped; a read keys as
  quality << 5 | reverse << 4 | base;
* standard_errmod (:406-488): keys are sorted, identical keys grouped, and
  the log10 likelihood table accumulated group by group in float, with
  quality clamped below at 4; diploid genotypes RR, RA, AA; haploid R, A (the
  haploid model adds its terms once per group, as the reference does).  Above
  max_depth keys the reference shuffles them, then sorts the whole list and
  reads only its first max_depth entries, so the shuffle has no effect: the
  max_depth smallest keys (lowest qualities) are used;
* glimpse_mpileup_reg (:236-249): with reads (depth > 0), GL = min(int(-10 *
  (llk - i)), 255) where i is the index of the largest llk (the reference
  subtracts the index, not the value), and the site stops being flat unless
  the values are equal.

The error tables are the reference's (genotype_bam_caller.h:36-42), generated
by gh_bam_tables in the library; this module generates them the same way and
tests/test_bam_caller.py pins entries of each against values read off the
reference header.
"""
import math

import numpy as np

f32 = np.float32
NT16 = "=ACMGRSVTWYHKDBN"
BASE4 = {"A": 0, "C": 1, "G": 2, "T": 3}
FUNMAP, FSECONDARY, FQCFAIL, FDUP = 0x4, 0x100, 0x200, 0x400
FPAIRED, FPROPER, FMUNMAP, FREVERSE, FMREVERSE = 0x1, 0x2, 0x8, 0x10, 0x20
DEFAULT = {"min_mapq": 10, "min_baseq": 10, "max_depth": 40, "fflag": FUNMAP | FSECONDARY | FQCFAIL | FDUP,
           "keep_orphan": 0, "check_orientation": 1, "check_proper_pair": 0}


def _g(x):
    return f32(float("%g" % x))


def tables():
    """log10(1 - e), log10(e / 3), log10(1/2 - e / 3), log10(e) for e = 10^(-q/10),
    printed to 6 significant digits, with the entries where the reference's
    printed table differs from that rule."""
    one_m = np.zeros(256, np.float32)
    div3 = np.zeros(256, np.float32)
    half = np.zeros(256, np.float32)
    err = np.zeros(256, np.float32)
    for q in range(256):
        e = 10.0 ** (-q / 10.0)
        one_m[q] = _g(math.log10(1.0 - e)) if 0 < q < 163 else (f32(0.99999) if q == 0 else f32(-4.82164e-17))
        div3[q] = _g(math.log10(e / 3.0))
        half[q] = _g(math.log10(0.5 - e / 3.0))
        err[q] = _g(-q / 10.0)
    one_m[61], one_m[97] = f32(-3.44972e-07), f32(-8.66531e-11)
    half[23], half[43] = f32(-0.302483), f32(-0.301044)
    return one_m, div3, half, err


def keep_read(r, prm):
    """read_bam's filter (genotype_bam_caller.cpp:30-59)."""
    fl = r["flag"]
    if r["tid"] < 0 or (fl & prm["fflag"]) or r["mapq"] < prm["min_mapq"]:
        return False
    if fl & FPAIRED:
        if fl & FMUNMAP:
            if not prm["keep_orphan"]:
                return False
        else:
            if prm["check_orientation"] and (
                    r["tid"] != r["mtid"] or ((fl & FREVERSE) and (fl & FMREVERSE)) or
                    (r["pos"] < r["mpos"] and (fl & FREVERSE)) or (r["pos"] > r["mpos"] and not (fl & FREVERSE))):
                return False
            if prm["check_proper_pair"] and not (fl & FPROPER):
                return False
    return True


def aligned_base(r, p):
    """(base index in the read, True) when reference position p (0-based) lies
    in an M / = / X operation of read r, else None."""
    ref, qp = r["pos"], 0
    for op, ln in r["cigar"]:
        if op in "M=X":
            if ref <= p < ref + ln:
                return qp + (p - ref)
            ref += ln
            qp += ln
        elif op in "IS":
            qp += ln
        elif op in "DN":
            if ref <= p < ref + ln:
                return None
            ref += ln
    return None


def call_site(keys, ref, alt, ploidy, prm, tabs):
    one_m, div3, half, err = tabs
    keys = sorted(keys)[:prm["max_depth"]]
    prr = pra = paa = f32(0.0)
    dp = 0
    j = 0
    while j < len(keys):
        b = keys[j]
        n = 1
        while j + n < len(keys) and keys[j + n] == b:
            n += 1
        q = max(b >> 5, 4)
        base = b & 0xF
        if ploidy > 1:
            em, d3, h = one_m[q], div3[q], half[q]
            prr = f32(prr + f32(em * f32(n))) if base == ref else f32(prr + f32(d3 * f32(n)))
            paa = f32(paa + f32(em * f32(n))) if base == alt else f32(paa + f32(d3 * f32(n)))
            pra = f32(pra + f32(h * f32(n))) if base in (ref, alt) else f32(pra + f32(d3 * f32(n)))
        else:
            prr = f32(prr + one_m[q]) if base == ref else f32(prr + err[q])
            paa = f32(paa + one_m[q]) if base == alt else f32(paa + err[q])
        dp += n
        j += n
    llk = [prr, pra, paa] if ploidy > 1 else [prr, paa]
    return llk, dp


def call_gls(reads, ploidy, sites, prm=None, beg=None, end=None):
    """reads: dicts tid, pos (0-based), mapq, flag, cigar [(op, len)], seq (str),
    qual (list), mtid, mpos; sites: dicts pos (1-based), ref, alt; all on one
    contig (tid 0).  Returns gl [L][ploidy+1], flat [L], depth [L]."""
    prm = dict(DEFAULT, **(prm or {}))
    tabs = tables()
    L = len(sites)
    P1 = ploidy + 1
    gl = np.zeros((L, P1), np.uint8)
    flat = np.ones(L, np.uint8)
    depth = np.zeros(L, np.int32)
    kept = [r for r in reads if keep_read(r, prm) and r["tid"] == 0]
    for l, s in enumerate(sites):
        p = s["pos"] - 1
        if beg is not None and s["pos"] < beg or end is not None and s["pos"] > end:
            continue
        if not (len(s["ref"]) == 1 and len(s["alt"]) == 1):
            continue   # indels: flat, no calling (the reference's default)
        ref, alt = BASE4.get(s["ref"].upper(), 4), BASE4.get(s["alt"].upper(), 4)
        if ref > 3 or alt > 3:
            raise ValueError("ERROR base sequences.")
        keys = []
        for r in kept:
            qp = aligned_base(r, p)
            if qp is None:
                continue
            b = BASE4.get(r["seq"][qp].upper(), 4)
            q = r["qual"][qp]
            if b > 3 or q < prm["min_baseq"]:
                continue
            keys.append(q << 5 | (1 if r["flag"] & FREVERSE else 0) << 4 | b)
        if not keys:
            continue
        llk, dp = call_site(keys, ref, alt, ploidy, prm, tabs)
        depth[l] = dp
        if dp > 0:
            mx = 0
            for j in range(1, P1):
                if llk[j] > llk[mx]:
                    mx = j
            for j in range(P1):
                v = f32(f32(-10.0) * f32(llk[j] - f32(mx)))
                gl[l, j] = min(int(v), 255)
            if not (gl[l, 0] == gl[l, 1] and gl[l, 0] == gl[l, ploidy]):
                flat[l] = 0
    return gl, flat, depth
