"""BAM genotype-likelihood caller (genotype_bam_caller.cpp, standard model).

CPU: the error tables (library and oracle) against entries read off the
reference's calling_utils tables (genotype_bam_caller.h:36-42); the oracle's
error model against a closed form of the likelihoods; its read filters on
hand-made reads.
GPU: gh_bam_call_gls against the oracle (oracle/bam_caller_ref.py) on
simulated BAM files (tests/bam_codec.py) -- GL, flat and depth bit-exact --
for diploid and haploid samples, filter settings, a region, and coverage
above max_depth."""
import math

import numpy as np
import pytest

import bam_caller_ref as br
import bam_codec as bcd

# (table, quality, value) as printed in the reference header
PINS = [(0, 1, -0.686825), (0, 4, -0.220481), (0, 10, -0.0457575), (0, 30, -0.000434512), (0, 61, -3.44972e-07),
        (0, 97, -8.66531e-11), (0, 163, -4.82164e-17), (0, 255, -4.82164e-17),
        (1, 0, -0.477121), (1, 10, -1.47712), (1, 96, -10.0771), (1, 255, -25.9771),
        (2, 0, -0.778151), (2, 1, -0.628519), (2, 23, -0.302483), (2, 43, -0.301044), (2, 58, -0.30103),
        (3, 0, 0.0), (3, 45, -4.5), (3, 255, -25.5)]


def test_oracle_tables_pinned():
    t = br.tables()
    for k, q, v in PINS:
        assert t[k][q] == np.float32(v), (k, q)


def test_library_tables_match_oracle():
    from glimpse_amd import bam
    t = bam.tables()
    want = br.tables()
    for k in range(4):
        assert np.array_equal(t[k], want[k]), k


def test_errmod_closed_form():
    """With one read per key (no grouping), the diploid llk are sums of
    log10 terms: RR gets log10(1-e) for ref bases, log10(e/3) otherwise; RA gets
    log10(1/2 - e/3) for ref or alt bases, log10(e/3) otherwise."""
    one_m, div3, half, _ = br.tables()
    keys = [30 << 5 | 0, 20 << 5 | 1 << 4 | 2, 12 << 5 | 1, 40 << 5 | 2]   # A(q30) G(q20,rev) C(q12) G(q40)
    llk, dp = br.call_site(keys, 0, 2, 2, dict(br.DEFAULT), br.tables())
    assert dp == 4
    order = sorted(keys)
    rr = np.float32(0)
    for b in order:
        q, base = b >> 5, b & 15
        rr = np.float32(rr + (one_m[q] if base == 0 else div3[q]))
    assert llk[0] == rr
    want_ra = sum(float(half[b >> 5]) if (b & 15) in (0, 2) else float(div3[b >> 5]) for b in keys)
    assert abs(float(llk[1]) - want_ra) < 1e-5
    # the max_depth rule keeps the smallest keys
    llk2, dp2 = br.call_site(keys, 0, 2, 2, dict(br.DEFAULT, max_depth=2), br.tables())
    assert dp2 == 2


def test_oracle_read_filters():
    base = {"tid": 0, "pos": 100, "mapq": 30, "flag": 0, "mtid": -1, "mpos": -1}
    p = dict(br.DEFAULT)
    assert br.keep_read(base, p)
    assert not br.keep_read(dict(base, mapq=9), p)
    assert not br.keep_read(dict(base, flag=0x400), p)
    assert not br.keep_read(dict(base, tid=-1), p)
    pr = dict(base, flag=0x1 | 0x20, mtid=0, mpos=300)           # forward read, mate downstream, reverse
    assert br.keep_read(pr, p)
    assert not br.keep_read(dict(pr, flag=0x1 | 0x10 | 0x20), p)  # both reverse
    assert not br.keep_read(dict(pr, mtid=1), p)
    assert not br.keep_read(dict(pr, flag=0x1 | 0x8), p)          # orphan
    assert br.keep_read(dict(pr, flag=0x1 | 0x8), dict(p, keep_orphan=1))
    assert not br.keep_read(pr, dict(p, check_proper_pair=1))
    assert br.keep_read(dict(pr, mtid=1), dict(p, check_orientation=0))


def _case(n_sites, n_reads, seed, L_ref=20000, indel_every=9):
    rng = np.random.default_rng(seed)
    pos = np.sort(rng.choice(np.arange(200, L_ref - 300), n_sites, replace=False)) + 1
    sites, ref, alt = [], [], []
    for i, p in enumerate(pos):
        r = "ACGT"[rng.integers(0, 4)]
        a = "ACGT".replace(r, "")[rng.integers(0, 3)]
        if i % indel_every == 4:
            ref.append(r + "T"), alt.append(r)      # a deletion: not called
        else:
            ref.append(r), alt.append(a)
            sites.append((int(p), r, a))
    haps = (rng.random((2, len(sites))) < 0.4).astype(int)
    reads = bcd.simulate(L_ref, sites, haps, n_reads, seed)
    # a second SNP record at an existing position (another ALT)
    k = 3
    pos = list(pos)
    pos.insert(k + 1, pos[k])
    ref.insert(k + 1, ref[k] if len(ref[k]) == 1 else "A")
    alt.insert(k + 1, "ACGT".replace(ref[k + 1], "")[0] if ref[k + 1] != alt[k] else alt[k])
    site_dicts = [{"pos": int(p), "ref": r, "alt": a} for p, r, a in zip(pos, ref, alt)]
    return reads, site_dicts, L_ref


def test_library_rejects_non_acgt_site(tmp_path):
    from glimpse_amd import bam, hip
    reads, sites, L_ref = _case(20, 50, 3)
    p = tmp_path / "x.bam"
    p.write_bytes(bcd.encode_bam([("chr1", L_ref)], reads))
    ref = [s["ref"] for s in sites]
    ref[0] = "N"
    with pytest.raises(hip.GhError, match="base sequences"):
        bam.call_gls(p, "chr1", [s["pos"] for s in sites], ref, [s["alt"] for s in sites])


GPU_CASES = [
    # name, ploidy, n_sites, n_reads, seed, params, region
    ("diploid", 2, 300, 3000, 1, {}, None),
    ("haploid", 1, 300, 3000, 2, {}, None),
    ("deep_max_depth", 2, 120, 12000, 3, {"max_depth": 12}, None),
    ("filters_off", 2, 200, 2500, 4, {"keep_orphan": 1, "check_orientation": 0, "min_mapq": 0, "min_baseq": 0}, None),
    ("proper_pairs_region", 2, 250, 2500, 5, {"check_proper_pair": 1, "fflag": 0x4}, (5000, 15000)),
]


@pytest.mark.gpu
@pytest.mark.parametrize("name,ploidy,n_sites,n_reads,seed,params,region", GPU_CASES, ids=[c[0] for c in GPU_CASES])
def test_bam_gls_match_oracle(tmp_path, name, ploidy, n_sites, n_reads, seed, params, region):
    from glimpse_amd import bam
    reads, sites, L_ref = _case(n_sites, n_reads, seed)
    p = tmp_path / "s.bam"
    p.write_bytes(bcd.encode_bam([("chr1", L_ref), ("chr2", 1000)], reads))
    beg, end = region if region else (0, 0)
    gl, flat, depth, st = bam.call_gls(p, "chr1", [s["pos"] for s in sites], [s["ref"] for s in sites],
                                       [s["alt"] for s in sites], ploidy=ploidy, beg=beg, end=end, **params)
    want_gl, want_flat, want_depth = br.call_gls(reads, ploidy, sites, params, beg=beg or None, end=end or None)
    assert np.array_equal(depth, want_depth)
    assert np.array_equal(gl, want_gl)
    assert np.array_equal(flat, want_flat)
    assert st["n_records"] == len(reads)
    called = int((flat == 0).sum())
    assert called > len(sites) // 4
    if name == "deep_max_depth":
        assert st["sites_over_max_depth"] > 0 and depth.max() == 12
