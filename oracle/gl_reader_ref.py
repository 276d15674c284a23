"""CPU restatement of the phase tool's target GL reading -- test
infrastructure only (the checker of gh_vcf_read_gls), never part of the
product path.

Follows phase/src/io/genotype_reader.cpp:352-521 (scanTarGenotypes +
parseTarGenotypes) for diploid targets, over records already parsed into
Python values: biallelic records only; the panel site is the first with the
record's position, REF and ALT (variant_map::getByRef); non-SNP records are
skipped unless indels are used; the record's field must hold ploidy + 1 = 3
values per sample (its largest count); per sample, a missing or vector-end
value skips the sample; PL is taken as is, GL as lroundf(-10 GL); a negative
PL / positive GL is an error; values are capped at 255; the site stays flat
for the sample only when its three values are equal.

Pinned in tests/test_vcf_io.py by hand-computed records (PL and GL, missing
and short vectors, indels, duplicate panel positions).
"""
import math

import numpy as np

MISSING = object()
END = object()


def _lroundf(x):
    x = float(np.float32(x))
    return int(math.floor(x + 0.5)) if x >= 0 else -int(math.floor(-x + 0.5))


def is_snp(ref, alt):
    return len(ref) == 1 and len(alt) == 1 and ref.upper() in "ACGTN" and alt.upper() in "ACGTN" and \
        ref.upper() != alt.upper()


def read_gls(records, n_samples, pos, ref, alt, use_gl=False, use_indels=False):
    """records: dicts with pos (1-based), alleles, and "PL" or "GL": per-sample
    lists (values, MISSING or END).  Returns gl [T][L][3], flat [T][L], n used."""
    L = len(pos)
    gl = np.zeros((n_samples, L, 3), np.uint8)
    flat = np.ones((n_samples, L), np.uint8)
    used = 0
    key = "GL" if use_gl else "PL"
    for r in records:
        if len(r["alleles"]) != 2:
            continue
        site = -1
        for l in range(L):
            if pos[l] == r["pos"] and ref[l] == r["alleles"][0] and alt[l] == r["alleles"][1]:
                site = l
                break
        if site < 0:
            continue
        if not use_indels and not is_snp(*r["alleles"]):
            continue
        vals = r.get(key)
        if vals is None or max(len(v) for v in vals) != 3:
            continue
        used += 1
        for i in range(n_samples):
            v = list(vals[i]) + [END] * (3 - len(vals[i]))
            if any(x is MISSING or x is END for x in v):
                continue
            if use_gl:
                llk = [_lroundf(-10 * np.float32(x)) for x in v]
                if min(llk) < 0:
                    raise ValueError("positive GL")
            else:
                llk = [int(x) for x in v]
                if min(llk) < 0:
                    raise ValueError("negative PL")
            g = [min(x, 255) for x in llk]
            gl[i, site] = g
            if not (g[0] == g[1] and g[0] == g[2]):
                flat[i, site] = 0
    return gl, flat, used
