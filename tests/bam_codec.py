"""BAM writer and read simulator for the BAM caller tests, written from the
SAM/BAM specification (SAM v1 section 4.2): BGZF blocks, "BAM\\1" header,
references, and records with CIGAR, 4-bit sequence and qualities."""
import struct

import numpy as np

from bcf_codec import bgzf_compress

CIGAR_OPS = "MIDNSHP=X"
NT16 = "=ACMGRSVTWYHKDBN"


def _reg2bin(beg, end):
    end -= 1
    if beg >> 14 == end >> 14:
        return ((1 << 15) - 1) // 7 + (beg >> 14)
    if beg >> 17 == end >> 17:
        return ((1 << 12) - 1) // 7 + (beg >> 17)
    if beg >> 20 == end >> 20:
        return ((1 << 9) - 1) // 7 + (beg >> 20)
    if beg >> 23 == end >> 23:
        return ((1 << 6) - 1) // 7 + (beg >> 23)
    if beg >> 26 == end >> 26:
        return ((1 << 3) - 1) // 7 + (beg >> 26)
    return 0


def ref_len(cigar):
    return sum(ln for op, ln in cigar if op in "MDN=X")


def encode_bam(refs, reads, text=None):
    """refs: [(name, length)]; reads: dicts name, tid, pos (0-based), mapq, flag,
    cigar [(op, len)], seq, qual (list of ints), mtid, mpos, tlen."""
    if text is None:
        text = "@HD\tVN:1.6\tSO:coordinate\n" + "".join("@SQ\tSN:%s\tLN:%d\n" % r for r in refs)
    t = text.encode()
    out = b"BAM\x01" + struct.pack("<i", len(t)) + t + struct.pack("<i", len(refs))
    for name, ln in refs:
        nb = name.encode() + b"\0"
        out += struct.pack("<i", len(nb)) + nb + struct.pack("<i", ln)
    for r in reads:
        name = r.get("name", "r").encode() + b"\0"
        cig = b"".join(struct.pack("<I", ln << 4 | CIGAR_OPS.index(op)) for op, ln in r["cigar"])
        seq = r["seq"]
        codes = [NT16.index(c) for c in seq]
        if len(codes) % 2:
            codes.append(0)
        sb = bytes(codes[i] << 4 | codes[i + 1] for i in range(0, len(codes), 2))
        qb = bytes(r["qual"])
        end = r["pos"] + max(ref_len(r["cigar"]), 1)
        bin_ = _reg2bin(r["pos"], end) if r["tid"] >= 0 else 4680
        body = struct.pack("<iiBBHHHiiii", r["tid"], r["pos"], len(name), r["mapq"], bin_, len(r["cigar"]),
                           r["flag"], len(seq), r.get("mtid", -1), r.get("mpos", -1), r.get("tlen", 0))
        body += name + cig + sb + qb
        out += struct.pack("<i", len(body)) + body
    return bgzf_compress(out)


def simulate(L_ref, sites, hap_alleles, n_reads, seed, read_len=(60, 150), junk=0.15):
    """Reads from a diploid sample over one contig of L_ref bases.  sites: list
    of (pos 1-based, ref, alt) SNPs; hap_alleles [2][n_sites] 0/1.  Reads carry
    the haplotype allele at covered sites, a random base elsewhere, base
    errors, varied CIGARs (soft clips, insertions, deletions, skips, =/X),
    qualities 2..60, both strands, and a share of flagged / low-MAPQ /
    odd-pair reads.  Returns reads sorted by position."""
    rng = np.random.default_rng(seed)
    genome = "".join(rng.choice(list("ACGT"), L_ref + 2000))   # slack for spans longer than a read
    genome = list(genome)
    for (p, ref, alt) in sites:
        genome[p - 1] = ref
    hap_seq = []
    for h in range(2):
        g = list(genome)
        for k, (p, ref, alt) in enumerate(sites):
            g[p - 1] = alt if hap_alleles[h][k] else ref
        hap_seq.append("".join(g))
    reads = []
    for i in range(n_reads):
        ln = int(rng.integers(*read_len))
        pos = int(rng.integers(0, L_ref - ln - 30))
        h = int(rng.integers(0, 2))
        src = hap_seq[h]
        cigar, seq, ref_at = [], [], pos
        # leading soft clip
        sc = int(rng.integers(0, 6)) if rng.random() < 0.3 else 0
        if sc:
            cigar.append(("S", sc))
            seq += list(rng.choice(list("ACGT"), sc))
        remaining = ln
        while remaining > 0:
            m = int(min(remaining, rng.integers(10, 60)))
            op = "M" if rng.random() < 0.8 else ("=" if rng.random() < 0.5 else "X")
            cigar.append((op, m))
            seq += list(src[ref_at:ref_at + m])
            ref_at += m
            remaining -= m
            if remaining <= 0:
                break
            u = rng.random()
            if u < 0.15:
                k = int(rng.integers(1, 4))
                cigar.append(("I", k))
                seq += list(rng.choice(list("ACGT"), k))
            elif u < 0.3:
                k = int(rng.integers(1, 4))
                cigar.append(("D", k))
                ref_at += k
            elif u < 0.35:
                k = int(rng.integers(5, 20))
                cigar.append(("N", k))
                ref_at += k
        if rng.random() < 0.2:
            k = int(rng.integers(1, 5))
            cigar.append(("S", k))
            seq += list(rng.choice(list("ACGT"), k))
        seq = "".join(seq)
        assert len(seq) == sum(ln for op, ln in cigar if op in "MIS=X")
        qual = [int(q) for q in rng.integers(2, 61, len(seq))]
        # base errors (and a few N)
        s = list(seq)
        for j in range(len(s)):
            u = rng.random()
            if u < 0.02:
                s[j] = rng.choice(list("ACGT"))
            elif u < 0.025:
                s[j] = "N"
        seq = "".join(s)
        rev = rng.random() < 0.5
        flag = 0x10 if rev else 0
        mapq = int(rng.integers(0, 61))
        mtid, mpos = -1, -1
        if rng.random() < 0.6:   # paired
            flag |= 0x1
            mtid, mpos = 0, pos + (int(rng.integers(50, 300)) if not rev else -int(rng.integers(50, 300)))
            mpos = max(mpos, 0)
            flag |= 0x20 if not rev else 0
            if rng.random() < 0.8:
                flag |= 0x2
            if rng.random() < junk / 3:
                flag |= 0x8          # mate unmapped
            if rng.random() < junk / 3:
                flag |= 0x20 if rev else 0   # both reverse
            if rng.random() < junk / 3:
                mtid = 1
        u = rng.random()
        if u < junk / 4:
            flag |= 0x400            # duplicate
        elif u < junk / 2:
            flag |= 0x100            # secondary
        elif u < 3 * junk / 4:
            flag |= 0x200            # QC fail
        reads.append({"name": "r%d" % i, "tid": 0, "pos": pos, "mapq": mapq, "flag": flag, "cigar": cigar,
                      "seq": seq, "qual": qual, "mtid": mtid, "mpos": mpos, "tlen": 0})
    # a few reads on another contig and unmapped ones
    reads.sort(key=lambda r: r["pos"])
    return reads
