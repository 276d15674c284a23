"""The phase tool's file-to-file path on the GPU: split_reference binary panel
-> target PLs from a VCF -> Gibbs run -> BCF output -> concordance against
the truth; and the same run from the panel's arrays gives identical output
(the binary panel decodes to exactly the arrays, stored sparse PBWT included)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_bin_vcf_run_bcf_concordance(tmp_path):
    from conftest import dosage_r2
    from glimpse_amd import binpanel, concordance, hip, synth, vcf
    P = synth.make_panel(1500, 1500, 4, length_mb=0.6, rare_frac=0.3, flip_frac=0.05, seed=31)
    gl, flat = synth.make_reads(P["tgt_haps"], 1.5, seed=7)
    L, T = P["L"], 4
    pos = np.arange(1, L + 1, dtype=np.int32) * 400 + 10000
    ref = ["ACGT"[i % 4] for i in range(L)]
    alt = ["CGTA"[i % 4] for i in range(L)]
    binf = tmp_path / "chunk.bin"
    binpanel.write(binf, P, pos, ref, alt, chrid="chr20")
    Q, info = binpanel.read(binf)
    assert info["n_sites"] == L
    # target PLs as a VCF (flat sites: no data)
    lines = ["##fileformat=VCFv4.2", "##contig=<ID=chr20>",
             '##FORMAT=<ID=PL,Number=G,Type=Integer,Description="PL">',
             "#CHROM\tPOS\tID\tREF\tALT\tQUAL\tFILTER\tINFO\tFORMAT\t" + "\t".join("t%d" % t for t in range(T))]
    for l in range(L):
        cells = ["." if flat[t, l] else "%d,%d,%d" % tuple(gl[t, l]) for t in range(T)]
        lines.append("chr20\t%d\t.\t%s\t%s\t.\t.\t.\tPL\t%s" % (pos[l], ref[l], alt[l], "\t".join(cells)))
    tf = tmp_path / "targets.vcf"
    tf.write_text("\n".join(lines) + "\n")
    samples, gl2, flat2, n = vcf.read_target_gls(tf, pos, ref, alt)
    assert n == int((flat == 0).any(axis=0).sum())   # records with no sample data hold one value: skipped
    data = flat == 0
    assert np.array_equal(gl2[data], gl[data]) and not flat2[data & ~(gl.min(axis=2) == gl.max(axis=2))].any()
    prm = hip.default_params(kpbwt=300, burnin=2, main_iter=4)
    a = hip.Panel(Q).phase_run(gl2, flat2, prm)
    b = hip.Panel(P).phase_run(gl2, flat2, prm)
    for k in ("gp0", "gp1", "hds", "h0", "h1", "nstates"):
        assert np.array_equal(a[k], b[k]), k
    f = hip.output_fields(a["gp0"], a["gp1"], a["hds"], P["cref"], P["calt"])
    out = tmp_path / "imputed.bcf"
    vcf.write_genotypes(out, "chr20", samples, pos, ref, alt, f, n_main=4)
    # read the BCF back and score it with the concordance kernels against the truth
    r = vcf.Reader(out)
    gp = np.zeros((L, T, 3), np.float32)
    ds = np.zeros((L, T), np.float32)
    for l, rec in enumerate(r):
        assert rec["pos"] == pos[l]
        gp[l] = r.format("GP", True)
        ds[l] = r.format("DS", True)[:, 0]
    truth = (P["tgt_haps"][0::2] + P["tgt_haps"][1::2]).astype(np.int8)   # [T][L]
    res = concordance.run(2, np.full(L, 0.3, np.float32), np.ones(L, np.uint8), esti_gp=gp, esti_ds=ds,
                          truth_gt=np.ascontiguousarray(truth.T), bins=[0, 0.5])
    s = concordance.summary(res, 2)
    assert s["n"] == T * L
    want = dosage_r2(a["gp0"], a["gp1"], truth)
    assert abs(s["r2_ds"] - want) < 2e-3    # DS written rounded to 1e-3
    assert s["r2_ds"] > 0.5
