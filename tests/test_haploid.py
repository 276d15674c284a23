"""Haploid targets (genotype ploidy 1; genotype.cpp:56-90, 127-166,
caller_algorithm.cpp:53-84, genotype_writer.cpp:138-190), on the CPU:
* the oracle's haploid run recovers a target that is a panel haplotype with
  certain GLs, and its GP / haplotype follow the stored posteriors;
* the oracle's output fields for haploid targets against closed forms (DS =
  gp1, two GP values on the 1e-3 grid summing to >= 0.9999, one allele, INFO 1
  unless every target is diploid, AF over the target haplotypes);
* the writer's haploid records decoded by the spec codec (vector ends)."""
import numpy as np
import pytest

import orc
from glimpse_amd import synth


def test_oracle_haploid_certain_gls():
    """A haploid target whose GLs are certain at every site (PL 0 / 255) is
    imputed to those alleles: its haplotype is exact, and GP is (1, 0) or
    (0, 1) except where no selected panel haplotype carries the allele in
    some main iteration (the target is a mosaic with mutations)."""
    P = synth.make_panel(300, 600, 2, length_mb=0.3, seed=7)
    L = P["L"]
    allele = P["tgt_haps"][0].astype(np.int32)
    gl = np.zeros((2, L, 3), np.uint8)
    gl[0, :, 0] = np.where(allele == 1, 255, 0)
    gl[0, :, 1] = np.where(allele == 1, 0, 255)
    gl[1] = synth.make_reads(P["tgt_haps"][2:4], 2.0, seed=1)[0][0]
    flat = np.zeros((2, L), np.uint8)
    prm = orc.default_params(kpbwt=100, burnin=1, main_iter=3, rng_mode=orc.RNG_COUNTER)
    out = orc.Panel(P).phase_run(gl, flat, prm, ploidy=[1, 2])
    assert np.array_equal(out["h0"][0], allele)
    assert np.all(out["h1"][0] == 0) and np.all(out["hds"][0] == 0)
    d = np.abs(out["gp1"][0] - allele)
    assert np.mean(d < 1e-3) > 0.97 and np.all(np.abs(out["gp0"][0] + out["gp1"][0] - 1) < 1e-5)


def test_oracle_haploid_output_fields_closed_forms():
    rng = np.random.default_rng(2)
    T, L = 5, 300
    a = rng.dirichlet([0.4, 0.4], (T, L)).astype(np.float32)
    pl = np.array([1, 2, 1, 2, 1], np.int32)
    gp0 = a[..., 0].copy()
    gp1 = a[..., 1].copy()
    d = rng.dirichlet([0.5, 0.5, 0.5], (T, L)).astype(np.float32)
    gp0[pl == 2], gp1[pl == 2] = d[pl == 2][..., 0], d[pl == 2][..., 1]
    cref, calt = np.full(L, 60, np.uint32), np.full(L, 40, np.uint32)
    o = orc.output_fields(gp0, gp1, np.zeros((T, L), np.uint8), cref, calt, ploidy=pl)
    h = pl == 1
    assert np.allclose(o["ds"][h], np.round(gp1[h].astype(np.float64), 3), atol=1.0001e-3)
    g = o["gp"][h]
    assert np.all(np.isnan(g[..., 2]))
    assert np.all(g[..., :2].sum(-1) >= 0.9999 - 1e-6) and np.all(np.abs(g[..., :2] - a[h]) <= 1.0011e-3)
    assert np.array_equal(o["gt"][h][..., 0], (gp0[h] < gp1[h]).astype(np.uint8)) and np.all(o["gt"][h][..., 1] == 255)
    assert np.all(o["info"] == 1.0)   # mixed ploidy: no INFO score
    ds_unr = np.where(h[:, None], gp1, gp1 + 2 * np.clip(1 - gp0 - gp1, 0, 1))
    assert np.allclose(o["af"], ds_unr.sum(0) / pl.sum(), atol=1e-5)
    # all haploid: INFO 1 as well; all diploid: the usual score (not all 1)
    oh = orc.output_fields(a[..., 0], a[..., 1], np.zeros((T, L), np.uint8), cref, calt, ploidy=np.ones(T, np.int32))
    assert np.all(oh["info"] == 1.0)
    od = orc.output_fields(d[..., 0], d[..., 1], np.zeros((T, L), np.uint8), cref, calt)
    assert not np.all(od["info"] == 1.0)


def test_writer_haploid_records(tmp_path):
    import bcf_codec as bc
    from glimpse_amd import vcf
    T, L = 3, 20
    rng = np.random.default_rng(4)
    gp = rng.dirichlet([1, 1, 1], (T, L)).astype(np.float32)
    gp[0, :, 2] = np.nan
    f = {"gt": rng.integers(0, 2, (T, L, 2)).astype(np.uint8), "gp": gp, "ds": gp[:, :, 1].copy(),
         "info": np.ones(L, np.float32), "af": rng.random(L).astype(np.float32), "raf": rng.random(L).astype(np.float32)}
    f["gt"][0, :, 1] = 255
    pos = np.arange(1, L + 1, dtype=np.int32) * 10
    p = tmp_path / "h.bcf"
    vcf.write_genotypes(p, "chrX", ["m", "f1", "f2"], pos, ["A"] * L, ["C"] * L, f, ploidy=[1, 2, 2])
    text, recs = bc.decode_bcf(p.read_bytes())
    assert "##FPLOIDY=-2" in text and "ID=INFO" not in text   # mixed ploidy: no INFO field
    for l, r in enumerate(recs):
        t, n, gts = r["fmt"][3]           # GT (PASS, RAF, AF, GT, DS, GP)
        assert gts[0] == [((int(f["gt"][0, l, 0]) + 1) << 1) | 1, -127]
        t, n, gps = r["fmt"][5]
        assert np.isnan(gps[0][2]) and np.array_equal(np.array(gps[0][:2], np.float32), gp[0, l, :2])
    import gzip
    assert gzip.decompress(p.read_bytes()).count(bytes.fromhex("0200807f")) >= L   # the float vector end
    v = tmp_path / "h.vcf"
    vcf.write_genotypes(v, "chrX", ["m", "f1", "f2"], pos, ["A"] * L, ["C"] * L, f, ploidy=[1, 2, 2])
    line = [x for x in v.read_text().splitlines() if not x.startswith("#")][0].split("\t")
    assert line[9].count(":") == 2 and "|" not in line[9].split(":")[0] and line[9].split(":")[2].count(",") == 1
