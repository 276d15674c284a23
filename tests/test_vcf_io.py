"""VCF / BCF input and output (host code of libglimpse_hip.so; no GPU):
* the target GL reader against the oracle restating genotype_reader.cpp:
  352-521, on VCF text, BGZF VCF, gzip VCF and BCF written by an independent
  Python codec (tests/bcf_codec.py), with hand-computed expectations for the
  rules (PL / GL, missing and short vectors, indels, duplicate positions);
* the writer (genotype_writer.cpp:43-190) decoded by that codec: header
  lines, BGZF framing and end-of-file block, every value bit for bit;
* writer -> reader round trips for BCF, BGZF VCF and VCF text.
htslib itself is not in the image, so byte compatibility with files htslib
writes is pinned only through the specification the codec follows."""
import gzip
import math
import struct

import numpy as np
import pytest

import bcf_codec as bc
import gl_reader_ref as gr

M, E = gr.MISSING, gr.END

PANEL = {  # pos, ref, alt
    "pos": [100, 200, 200, 300, 400, 500, 600, 700],
    "ref": ["A", "C", "C", "G", "T", "AT", "G", "C"],
    "alt": ["G", "T", "T", "A", "C", "A", "T", "A"],
}


def _records():
    """Target records: (pos, alleles, PL per sample, GL per sample) for 3 samples."""
    return [
        {"pos": 100, "alleles": ["A", "G"], "PL": [[0, 10, 100], [0, 0, 0], [M, 3, 4]],
         "GL": [[0.0, -1.0, -10.0], [0.0, 0.0, 0.0], [M, -0.3, -0.4]]},
        {"pos": 150, "alleles": ["A", "C"], "PL": [[0, 5, 9]] * 3, "GL": [[0.0, -0.5, -0.9]] * 3},   # not in panel
        {"pos": 200, "alleles": ["C", "T"], "PL": [[300, 0, 20], [7, 7, 7], [1, 2, 3]],               # duplicate panel pos
         "GL": [[-30.04, 0.0, -2.0], [-0.7, -0.7, -0.7], [-0.14, -0.25, -0.35]]},
        {"pos": 300, "alleles": ["G", "A", "C"], "PL": [[0, 1, 2, 3, 4, 5]] * 3,                       # multi-allelic
         "GL": [[0.0, -0.1, -0.2, -0.3, -0.4, -0.5]] * 3},
        {"pos": 400, "alleles": ["T", "C"], "PL": [[0, 30], [0, 30], [0, 30]],                          # 2 values
         "GL": [[0.0, -3.0]] * 3},
        {"pos": 500, "alleles": ["AT", "A"], "PL": [[0, 20, 40]] * 3, "GL": [[0.0, -2.0, -4.0]] * 3},  # indel
        {"pos": 600, "alleles": ["G", "T"], "PL": [[12, 0, 255], [0, E, E], [40, 20, 0]],             # one short sample
         "GL": [[-1.2, 0.0, -25.5], [0.0, E, E], [-4.0, -2.0, 0.0]]},
        {"pos": 700, "alleles": ["C", "G"], "PL": [[0, 1, 2]] * 3, "GL": [[0.0, -0.1, -0.2]] * 3},    # ALT differs
    ]


def _vcf_text(records, key):
    samples = ["s0", "s1", "s2"]
    h = ["##fileformat=VCFv4.2", "##contig=<ID=chr20>",
         '##FORMAT=<ID=PL,Number=G,Type=Integer,Description="PL">',
         '##FORMAT=<ID=GL,Number=G,Type=Float,Description="GL">',
         '##FORMAT=<ID=DP,Number=1,Type=Integer,Description="DP">',
         "#CHROM\tPOS\tID\tREF\tALT\tQUAL\tFILTER\tINFO\tFORMAT\t" + "\t".join(samples)]
    lines = []
    for r in records:
        cells = []
        for v in r[key]:
            cells.append(",".join("." if x is M else str(x) for x in v if x is not E) + ":5")
        lines.append("\t".join(["chr20", str(r["pos"]), ".", r["alleles"][0], ",".join(r["alleles"][1:]), ".",
                                "PASS", ".", key + ":DP"] + cells))
    return "\n".join(h + lines) + "\n", samples


def _bcf_bytes(records, key):
    text, samples = _vcf_text([], key)
    text = text.replace("##contig", '##FILTER=<ID=PASS,Description="All filters passed">\n##contig')
    dictionary = {"PASS": 0, "PL": 1, "GL": 2, "DP": 3}
    recs = []
    for r in records:
        kind = "f" if key == "GL" else "i16"
        per = [[None if x is M else x for x in v if x is not E] for v in r[key]]
        recs.append({"chrom": "chr20", "pos": r["pos"], "alleles": r["alleles"],
                     "fmt": [(key, kind, per), ("DP", "i8", [[5]] * 3)]})
    return bc.encode_bcf(text, samples, recs, dictionary, ["chr20"])


def _want(key, use_indels=False):
    return gr.read_gls(_records(), 3, PANEL["pos"], PANEL["ref"], PANEL["alt"], use_gl=key == "GL",
                       use_indels=use_indels)


def test_oracle_hand_cases():
    gl, flat, n = _want("PL")
    L = {p: i for i, p in reversed(list(enumerate(PANEL["pos"])))}
    assert n == 3                                              # pos 100, 200, 600
    assert gl[0, L[100]].tolist() == [0, 10, 100] and flat[0, L[100]] == 0
    assert gl[1, L[100]].tolist() == [0, 0, 0] and flat[1, L[100]] == 1
    assert gl[2, L[100]].tolist() == [0, 0, 0] and flat[2, L[100]] == 1   # missing value: skipped
    assert gl[0, 1].tolist() == [255, 0, 20] and gl[:, 2].sum() == 0     # first of the duplicate panel sites
    assert gl[0, L[600]].tolist() == [12, 0, 255] and flat[1, L[600]] == 1  # short vector: skipped
    assert gl[:, L[500]].sum() == 0 and flat[:, L[500]].all()            # indel skipped
    gi, _, ni = _want("PL", use_indels=True)
    assert ni == 4 and gi[0, L[500]].tolist() == [0, 20, 40]
    g2, f2, n2 = _want("GL")
    assert g2[0, L[100]].tolist() == [0, 10, 100]
    assert g2[0, 1].tolist() == [255, 0, 20]                   # lroundf(300.4) capped at 255
    assert g2[2, 1].tolist() == [1, 3, 4]                      # lroundf: 1.4 -> 1, 2.5 -> 3 (half away from 0)
    assert f2[1, 1] == 1


def _check_reader(path, key, use_indels=False):
    from glimpse_amd import vcf
    samples, gl, flat, n = vcf.read_target_gls(path, PANEL["pos"], PANEL["ref"], PANEL["alt"], use_gl=key == "GL",
                                               use_indels=use_indels)
    wg, wf, wn = _want(key, use_indels)
    assert samples == ["s0", "s1", "s2"]
    assert n == wn
    assert np.array_equal(gl, wg) and np.array_equal(flat, wf)


@pytest.mark.parametrize("key", ["PL", "GL"])
@pytest.mark.parametrize("form", ["vcf", "vcf.gz", "gzip", "bcf"])
def test_read_target_gls(tmp_path, key, form):
    text, _ = _vcf_text(_records(), key)
    p = tmp_path / ("t." + form)
    if form == "vcf":
        p.write_text(text)
    elif form == "vcf.gz":
        p.write_bytes(bc.bgzf_compress(text.encode()))
    elif form == "gzip":
        p.write_bytes(gzip.compress(text.encode()))
    else:
        p.write_bytes(_bcf_bytes(_records(), key))
    _check_reader(p, key)
    _check_reader(p, key, use_indels=True)


def test_negative_pl_is_an_error(tmp_path):
    from glimpse_amd import hip, vcf
    recs = [{"pos": 100, "alleles": ["A", "G"], "PL": [[0, -3, 10], [0, 0, 0], [0, 0, 0]]}]
    text, _ = _vcf_text(recs, "PL")
    p = tmp_path / "neg.vcf"
    p.write_text(text)
    with pytest.raises(hip.GhError, match="negative value in FORMAT/PL"):
        vcf.read_target_gls(p, PANEL["pos"], PANEL["ref"], PANEL["alt"])


def test_record_reader_values(tmp_path):
    """The record reader hands out htslib's values: padding with the vector
    end, the missing sentinel, GT codes."""
    from glimpse_amd import vcf
    p = tmp_path / "r.vcf"
    p.write_text("\n".join([
        "##fileformat=VCFv4.2", "##contig=<ID=1>",
        '##INFO=<ID=AF,Number=A,Type=Float,Description="x">', '##INFO=<ID=AC,Number=A,Type=Integer,Description="x">',
        '##FORMAT=<ID=GT,Number=1,Type=String,Description="x">', '##FORMAT=<ID=PL,Number=G,Type=Integer,Description="x">',
        '##FORMAT=<ID=DS,Number=1,Type=Float,Description="x">',
        "#CHROM\tPOS\tID\tREF\tALT\tQUAL\tFILTER\tINFO\tFORMAT\ta\tb\tc",
        "1\t10\trs1\tA\tT\t.\tPASS\tAF=0.25;AC=3\tGT:PL:DS\t0|1:0,3,6:1.5\t1/1:.:.\t.:4,5:0",
        "1\t11\t.\tA\tT,C\t.\tPASS\t.\tGT\t0\t.|1\t1|2"]) + "\n")
    r = vcf.Reader(p)
    it = iter(r)
    rec = next(it)
    assert (rec["chrom"], rec["pos"], rec["id"], rec["n_allele"], rec["ref"], rec["alt"]) == ("1", 10, "rs1", 2, "A", "T")
    assert rec["is_snp"]
    assert r.info("AF", True).tolist() == [0.25] and r.info("AC").tolist() == [3]
    assert r.format("GT").tolist() == [[2, 5], [4, 4], [0, vcf.I32_END]]
    assert r.format("PL").tolist() == [[0, 3, 6], [vcf.I32_MISSING, vcf.I32_END, vcf.I32_END], [4, 5, vcf.I32_END]]
    ds = r.format("DS", True)
    assert ds[0, 0] == 1.5 and ds.view(np.uint32)[1, 0] == bc.F32_MISSING
    rec = next(it)
    assert rec["n_allele"] == 3 and rec["id"] == "."
    assert r.format("GT").tolist() == [[2, vcf.I32_END], [0, 5], [4, 7]]
    assert r.format("PL").size == 0
    with pytest.raises(StopIteration):
        next(it)


def _fields(T, L, seed):
    rng = np.random.default_rng(seed)
    gp = rng.random((T, L, 3)).astype(np.float32)
    gp /= gp.sum(axis=2, keepdims=True)
    gp = (np.floor(gp * 1000) / 1000).astype(np.float32)
    return {"gt": rng.integers(0, 2, (T, L, 2)).astype(np.uint8), "gp": gp,
            "ds": (gp[:, :, 1] + 2 * gp[:, :, 2]).astype(np.float32),
            "info": rng.random(L).astype(np.float32), "af": rng.random(L).astype(np.float32),
            "raf": rng.random(L).astype(np.float32)}


def _sites(L):
    pos = np.arange(1, L + 1, dtype=np.int32) * 37
    ref = ["ACGT"[i % 4] for i in range(L)]
    alt = ["CGTA"[i % 4] if i % 7 else "CGTAA"[: 1 + i % 3] for i in range(L)]
    ids = ["rs%d" % i if i % 3 else "." for i in range(L)]
    return pos, ref, alt, ids


def test_writer_bcf_decoded_by_spec_codec(tmp_path):
    from glimpse_amd import vcf
    T, L = 5, 9000   # several BGZF blocks
    f = _fields(T, L, 1)
    pos, ref, alt, ids = _sites(L)
    p = tmp_path / "o.bcf"
    vcf.write_genotypes(p, "chr20", ["n%d" % i for i in range(T)], pos, ref, alt, f, ids=ids, n_main=15,
                        file_date="20260101")
    raw = p.read_bytes()
    blocks = bc.bgzf_blocks(raw)
    assert len(blocks) > 2 and raw.endswith(bc.BGZF_EOF) and blocks[-1] == (28, 0)
    assert all(isize <= 0xFF00 for _, isize in blocks)
    text, recs = bc.decode_bcf(raw)
    lines = text.splitlines()
    for want in ["##fileformat=VCFv4.2", "##fileDate=20260101", "##contig=<ID=chr20>", "##NMAIN=15", "##FPLOIDY=2",
                 '##FORMAT=<ID=GP,Number=3,Type=Float,Description="Genotype posteriors">']:
        assert want in lines
    assert lines[-1] == "#CHROM\tPOS\tID\tREF\tALT\tQUAL\tFILTER\tINFO\tFORMAT\t" + "\t".join("n%d" % i for i in range(T))
    assert len(recs) == L
    f32 = lambda a: [struct.unpack("<f", struct.pack("<f", float(x)))[0] for x in a]
    for l in (0, 1, 6, 7, L - 1):
        r = recs[l]
        assert (r["rid"], r["pos"], r["rlen"], r["n_allele"], r["n_sample"], r["n_fmt"]) == (0, int(pos[l]), len(ref[l]), 2, T, 3)
        assert r["qual_bits"] == bc.F32_MISSING and r["filter"] == []
        assert r["id"] == ("" if ids[l] == "." else ids[l]) and r["alleles"] == [ref[l], alt[l]]
        assert r["info"] == {1: f32([f["raf"][l]]), 2: f32([f["af"][l]]), 3: f32([f["info"][l]])}
        t, n, gts = r["fmt"][4]
        assert (t, n) == (1, 2)
        assert gts == [[((int(f["gt"][i, l, 0]) + 1) << 1) | 1, ((int(f["gt"][i, l, 1]) + 1) << 1) | 1] for i in range(T)]
        assert r["fmt"][5] == (5, 1, [f32([f["ds"][i, l]]) for i in range(T)])
        assert r["fmt"][6] == (5, 3, [f32(f["gp"][i, l]) for i in range(T)])


@pytest.mark.parametrize("ext", ["bcf", "vcf.gz", "vcf"])
def test_writer_reader_round_trip(tmp_path, ext):
    from glimpse_amd import vcf
    T, L = 4, 700
    f = _fields(T, L, 2)
    pos, ref, alt, ids = _sites(L)
    p = tmp_path / ("o." + ext)
    vcf.write_genotypes(p, "chr7", ["a", "b", "c", "d"], pos, ref, alt, f, ids=ids)
    if ext == "vcf.gz":
        assert p.read_bytes().endswith(bc.BGZF_EOF)
        bc.bgzf_blocks(p.read_bytes())
    r = vcf.Reader(p)
    assert r.samples == ["a", "b", "c", "d"]
    n = 0
    for l, rec in enumerate(r):
        assert (rec["chrom"], rec["pos"], rec["ref"], rec["alt"]) == ("chr7", int(pos[l]), ref[l], alt[l])
        assert rec["id"] == ids[l]
        gt = r.format("GT")
        a = (gt >> 1) - 1
        assert np.array_equal(a, f["gt"][:, l, :].astype(np.int32))
        assert np.all(gt[:, 1] & 1)
        ds, gp = r.format("DS", True)[:, 0], r.format("GP", True)
        raf = r.info("RAF", True)[0]
        if ext == "bcf":
            assert np.array_equal(ds, f["ds"][:, l]) and np.array_equal(gp, f["gp"][:, l]) and raf == f["raf"][l]
        else:   # text: %g, six significant digits
            assert np.allclose(ds, f["ds"][:, l], rtol=1e-5, atol=1e-6) and np.allclose(gp, f["gp"][:, l], rtol=1e-5, atol=1e-6)
            assert math.isclose(raf, f["raf"][l], rel_tol=1e-5, abs_tol=1e-6)
        n += 1
    assert n == L
